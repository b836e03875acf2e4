// Error state and version of the C ABI.
#include <string>

#include "spmv_amd.h"

namespace spmvx
{
namespace
{
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
} // namespace spmvx

extern "C" {
const char* spmvx_last_error(void) { return spmvx::g_last_error.c_str(); }
int spmvx_abi_version(void) { return SPMVX_ABI_VERSION; }
}
