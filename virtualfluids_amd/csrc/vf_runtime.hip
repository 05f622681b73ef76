// Library-wide bookkeeping: ABI version and the last error message.
#include <cstdio>
#include <cstring>

#include "vf_common.h"

namespace vfa {

static thread_local char g_last_error[512] = "";

void set_last_error(const char* msg)
{
    std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg);
}

int check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
        return VF_ERR_HIP;
    }
    return VF_OK;
}

} // namespace vfa

extern "C" int vf_abi_version(void) { return VF_ABI_VERSION; }

extern "C" const char* vf_get_last_error(void) { return vfa::g_last_error; }
