// amx_common.cpp -- last-error text, version and device queries of the C ABI.
#include "amx_internal.hpp"

namespace amx {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
}  // namespace amx

extern "C" {

size_t amx_last_error(char* buf, size_t len)
{
    const std::string& m = amx::g_last_error;
    if (buf && len > 0) {
        const size_t n = m.size() < len - 1 ? m.size() : len - 1;
        std::memcpy(buf, m.data(), n);
        buf[n] = '\0';
    }
    return m.size();
}

const char* amx_version(void) { return "amat_amd 0.1.0 gfx950"; }

int amx_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // extern "C"
