// obme_internal.h — helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/obme.h"

// Records `msg` as this thread's last error (obme_last_error) and returns code.
obme_status obme_fail(obme_status code, const std::string& msg);

#define OBME_HIP_TRY(expr)                                                                      \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return obme_fail(OBME_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)
