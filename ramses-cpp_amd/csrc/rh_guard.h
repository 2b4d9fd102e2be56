// rh_guard.h -- no C++ exception crosses the C-ABI.
//
// The entry points build host containers (level tables, sort buffers, tree
// lists); std::bad_alloc, std::length_error or any other exception thrown
// there would otherwise leave an extern "C" function and end the host
// process in std::terminate.  Every int-returning entry point runs its body
// through guard(), which turns an exception into RH_ERR_HOST with its text
// in the error string(s) the caller reads back (rh_last_error).
#pragma once
#include <exception>
#include <new>
#include <string>

#include "rhydro.h"

namespace rhi {

// Error text of calls that have no context (rh_create, the driver helpers).
extern std::string g_create_error;

inline void set_error_text(std::string* e, const char* msg) noexcept {
    if (!e) return;
    try {
        *e = msg;
    } catch (...) {
        // The message itself could not be stored (memory exhausted): the
        // status code still reports the failure.
    }
}

// Runs body() and returns its status; an exception becomes RH_ERR_HOST with
// its message stored in *e1 and *e2 (either may be null).
template <class F> int guard(std::string* e1, std::string* e2, F&& body) noexcept {
    const char* msg = "unknown C++ exception";
    try {
        return body();
    } catch (const std::bad_alloc&) {
        msg = "host memory exhausted (std::bad_alloc)";
    } catch (const std::exception& ex) {
        msg = ex.what();
    } catch (...) {
    }
    set_error_text(e1, msg);
    set_error_text(e2, msg);
    return RH_ERR_HOST;
}

}  // namespace rhi
