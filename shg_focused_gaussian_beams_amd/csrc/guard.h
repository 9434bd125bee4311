// guard.h — keeps C++ exceptions from crossing the C ABI.  Every `int` entry point of
// include/shg.h runs its body through shg_internal_guard: a failed host allocation
// becomes SHG_E_NOMEM, any other exception SHG_E_INTERNAL, each with a message for
// shg_last_error(), instead of unwinding into a C caller.
#pragma once
#include <exception>
#include <new>
#include <string>

#include "../../include/shg.h"

int shg_internal_fail(int code, const std::string &msg);

template <class F>
int shg_internal_guard(const char *fn, F &&body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return shg_internal_fail(SHG_E_NOMEM, std::string(fn) + ": host allocation failed");
    } catch (const std::exception &e) {
        return shg_internal_fail(SHG_E_INTERNAL, std::string(fn) + ": " + e.what());
    } catch (...) {
        return shg_internal_fail(SHG_E_INTERNAL, std::string(fn) + ": unexpected error");
    }
}
