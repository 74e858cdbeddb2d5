// capi.hip -- error reporting and shared host-side constants of libholohip.
#include "common.hpp"
#include "../../include/holohip.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <vector>

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

// ---- optional launch timing (HIP events on the launch stream) ----
struct TimedLaunch {
    const char* name;   // static string
    hipEvent_t start, stop;
};
static std::mutex g_tmu;
static bool g_timing = false;
static std::vector<TimedLaunch> g_timed;

KTimer::KTimer(const char* name, hipStream_t st) : name_(name), st_(st), on_(false) {
    std::lock_guard<std::mutex> lk(g_tmu);
    if (!g_timing) return;
    if (hipEventCreate(&a_) != hipSuccess) return;
    if (hipEventCreate(&b_) != hipSuccess) {
        (void)hipEventDestroy(a_);
        return;
    }
    on_ = hipEventRecord(a_, st_) == hipSuccess;
}

KTimer::~KTimer() {
    if (!on_) return;
    (void)hipEventRecord(b_, st_);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timed.push_back(TimedLaunch{name_, a_, b_});
}

static void timing_clear_locked() {
    for (auto& t : g_timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    g_timed.clear();
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

extern "C" int hd_timing_enable(int on) {
    std::lock_guard<std::mutex> lk(hd::g_tmu);
    hd::timing_clear_locked();
    hd::g_timing = on != 0;
    return HD_OK;
}

extern "C" int hd_timing_read(const char* name, double* total_ms, int* launches) {
    if (!name || !total_ms || !launches) return hd::set_error(HD_ERR_ARG, "hd_timing_read: null argument");
    std::lock_guard<std::mutex> lk(hd::g_tmu);
    double tot = 0.0;
    int n = 0;
    for (auto& t : hd::g_timed) {
        if (strcmp(t.name, name) != 0) continue;
        if (hipEventSynchronize(t.stop) != hipSuccess) return hd::check_launch("hd_timing_read");
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) != hipSuccess) return hd::check_launch("hd_timing_read");
        tot += ms;
        n++;
    }
    *total_ms = tot;
    *launches = n;
    return HD_OK;
}
