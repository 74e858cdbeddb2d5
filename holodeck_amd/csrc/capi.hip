// capi.hip -- error reporting and shared host-side constants of libholohip.
#include "common.hpp"
#include "../../include/holohip.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>

namespace hd {

static thread_local char g_msg[512] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(HD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return HD_OK;
}

// sam_cyutils.pyx:36-41, same expression order as the module initialiser.
EvoConsts evo_consts() {
    EvoConsts K;
    K.gw_dadt_sep = -64.0 * pow(MY_NWTG, 3) / 5.0 / pow(MY_SPLC, 5);
    K.kepler_freq = (1.0 / (2.0 * M_PI)) * sqrt(MY_NWTG);
    K.kepler_sepa = pow(MY_NWTG, 1.0 / 3.0) / pow(2.0 * M_PI, 2.0 / 3.0);
    K.four_pi_c_mpc = 4 * M_PI * MY_SPLC / MY_MPC;
    return K;
}

}  // namespace hd

extern "C" const char* hd_last_error(void) { return hd::g_msg; }
extern "C" int hd_version(void) { return 1; }
