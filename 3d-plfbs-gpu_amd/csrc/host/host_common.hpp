// Shared host helpers: error reporting for the C-ABI.
#pragma once
#include <cstdarg>
#include <cstdint>
#include <new>
#include <stdexcept>

#include "plbfs_hip.h"

// Records a formatted message as the last error and returns rc.
int plbfs_fail(int rc, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// Runs fn() and turns an escaping allocation failure or library exception into
// PLBFS_ERR_INPUT with the reason as the last error, so that no exception
// crosses the C-ABI.
template <class Fn>
int plbfs_guarded(const char *what, Fn fn) {
    try {
        return fn();
    } catch (const std::bad_alloc &) {
        return plbfs_fail(PLBFS_ERR_INPUT, "%s: out of host memory", what);
    } catch (const std::exception &e) {
        return plbfs_fail(PLBFS_ERR_INPUT, "%s: %s", what, e.what());
    }
}
