// C-ABI plumbing: thread-local error text and device information.
#include <string.h>

#include "qdfn_internal.h"

namespace qdfn {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace qdfn

extern "C" const char* qdfn_last_error(void) { return qdfn::g_last_error.c_str(); }

extern "C" int qdfn_device_info(char* name, int64_t name_len, int32_t* cu_count, int64_t* hbm_bytes) {
    int dev = 0;
    QDFN_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    QDFN_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len > 0) {
        strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
    return QDFN_OK;
}
