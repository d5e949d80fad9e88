// lem_capi.hip -- C-ABI plumbing: version, thread-local error text, launch checks.
//
// The reference reports failures with std::runtime_error through pybind11 and
// aborts the interpreter with exit() on CUDA errors (gpu_error.cuh:2-10).  This
// library never exits: every entry point returns 0 or a negative LEM_E_* code
// and leaves a message for lem_last_error().
#include <hip/hip_runtime.h>

#include <string>

#include "lem_internal.h"
#include "../../include/lem.h"

namespace lem {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return E_HIP;
    }
    return OK;
}

}  // namespace lem

extern "C" int lem_abi_version(void) { return LEM_ABI_VERSION; }

extern "C" const char *lem_last_error(void) { return lem::g_last_error.c_str(); }

extern "C" int lem_device_arch(char *buf, int len)
{
    lem::clear_error();
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        lem::set_error("lem_device_arch: no HIP device");
        return lem::E_HIP;
    }
    if (buf && len > 0) {
        int i = 0;
        for (; i < len - 1 && prop.gcnArchName[i]; ++i)
            buf[i] = prop.gcnArchName[i];
        buf[i] = 0;
    }
    return lem::OK;
}
