// Error reporting, device selection and version of the ed_hip C-ABI.
#include <cstring>
#include <string>

#include "ed_internal.h"

namespace ed {

static thread_local std::string g_last_error = "no error";

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

}  // namespace ed

extern "C" {

const char *ed_last_error(void) { return ed::g_last_error.c_str(); }

const char *ed_version(void) { return "ed_hip 0.1 gfx950"; }

int ed_init(int device)
{
    int n = 0;
    ED_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return ed::fail(ED_ERR_ARG, "ed_init: device " + std::to_string(device) + " out of range");
    ED_HIP_CHECK(hipSetDevice(device));
    return ed::noalloc_setup(device);
}

int ed_lanczos_default_opts(ed_lanczos_opts *o)
{
    if (!o) return ed::fail(ED_ERR_ARG, "ed_lanczos_default_opts: NULL");
    o->nitermax = 128;      // Nitermax, dmft_ed_solver_variables.h:18
    o->neigen = 1;          // Neigen
    o->tolerance = 1e-12;   // tolerance, dmft_ed_solver_variables.h:20
    o->precision = ED_PREC_F64;
    o->check_every = 8;
    o->reorth = 0;
    return ED_OK;
}

}  // extern "C"
