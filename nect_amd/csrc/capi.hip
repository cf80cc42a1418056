// Error reporting shared by every C-ABI entry point.
#include "common.hpp"

namespace nect {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int32_t check_launch(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return 1;
    }
    return 0;
}

}  // namespace nect

extern "C" const char *nect_last_error(void) { return nect::g_last_error.c_str(); }

extern "C" int32_t nect_abi_version(void) { return NECT_ABI_VERSION; }
