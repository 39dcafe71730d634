// Shared host helpers: error reporting for the C-ABI.
#pragma once
#include <cstdarg>
#include <cstdint>

#include "plbfs_hip.h"

// Records a formatted message as the last error and returns rc.
int plbfs_fail(int rc, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
