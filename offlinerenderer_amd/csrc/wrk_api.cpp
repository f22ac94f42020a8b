// wrk_api.cpp — error state and device queries of the wrk_* C ABI.
#include <hip/hip_runtime.h>

#include <string>

#include "wrk_common.h"

namespace wrk {
static thread_local std::string g_lastError;
void set_error(const std::string& msg) { g_lastError = msg; }
}  // namespace wrk

extern "C" const char* wrk_last_error(void) { return wrk::g_lastError.c_str(); }

extern "C" int wrk_device_count(int* count) {
    WRK_REQUIRE(count, "wrk_device_count: null count");
    WRK_HIP_TRY(hipGetDeviceCount(count));
    return WRK_OK;
}
