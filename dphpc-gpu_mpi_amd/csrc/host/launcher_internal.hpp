// Pieces of the host launcher shared by its two launch paths (launcher.cpp:
// the world in device allocations; noalloc.hip: the world in a static
// device array).  Internal to libgpumpi.so.
#pragma once

#include "gpumpi/host.h"
#include "gpumpi_rt/world.hpp"

#define GPUMPI_HIPOK(call)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            ::gpumpi::host::set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),     \
                                    __FILE__, __LINE__);                                       \
            return GPUMPI_EHIP;                                                                \
        }                                                                                      \
    } while (0)

namespace gpumpi {
namespace host {

extern thread_local char g_err[512];
void set_err(const char* fmt, ...);
// Checked launch configuration and the rank workgroup's LDS bytes.
int build_config(const gpumpi_config* in, rt::Config* out, int* lds_bytes);
// Host service channel (service.cpp): slots for `nranks` ranks, all idle,
// served by svc_poll until the next svc_prepare.
int svc_prepare(int nranks, rt::SvcSlot** out);
int svc_poll();

}  // namespace host
}  // namespace gpumpi
