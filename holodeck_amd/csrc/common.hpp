// common.hpp -- shared device helpers, constants and C-ABI plumbing for libholohip.
//
// Constants follow holodeck/sams/sam_cyutils.pyx:30-41 and holodeck/cyutils.pyx:39-48
// (the Cython modules carry their own truncated copies of G and Mpc; the
// Python-side astropy values are passed in from the host where the reference
// uses them).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define HD_OK 0
#define HD_ERR_ARG -1
#define HD_ERR_HIP -2
#define HD_ERR_SIZE -3

namespace hd {

// ---- error reporting (thread-local message, int status codes) ----
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

// Records HIP events around a launch sequence on its stream while timing is
// enabled (hd_timing_enable); a no-op otherwise.
class KTimer {
  public:
    KTimer(const char* name, hipStream_t st);
    ~KTimer();
    KTimer(const KTimer&) = delete;
    KTimer& operator=(const KTimer&) = delete;

  private:
    const char* name_;
    hipStream_t st_;
    hipEvent_t a_, b_;
    bool on_;
};

// ---- sam_cyutils.pyx:30-41 ----
constexpr double MY_NWTG = 6.6742999e-08;
constexpr double MY_SPLC = 29979245800.0;
constexpr double MY_MPC = 3.08567758e+24;
constexpr double MY_SCHW = 1.4852320538237328e-28;

// Derived constants are computed once on the host with the same expression
// order as the Cython module initialiser and passed by value to kernels.
struct EvoConsts {
    double gw_dadt_sep;       // -64 G^3 / 5 / c^5
    double kepler_freq;       // sqrt(G) / (2 pi)
    double kepler_sepa;       // G^(1/3) / (2 pi)^(2/3)
    double four_pi_c_mpc;     // 4 pi c / Mpc
};
EvoConsts evo_consts();

// sam_cyutils.pyx:45-49
__device__ __forceinline__ double hard_gw(const EvoConsts& K, double mtot, double mrat, double sepa) {
    double m3 = mtot * mtot * mtot;
    double a3 = sepa * sepa * sepa;
    double q1 = 1.0 + mrat;
    return K.gw_dadt_sep * m3 * mrat / a3 / (q1 * q1);
}

// sam_cyutils.pyx:52-61
__device__ __forceinline__ double kepler_freq_from_sepa(const EvoConsts& K, double mtot, double sepa) {
    return K.kepler_freq * sqrt(mtot) / (sepa * sqrt(sepa));
}
__device__ __forceinline__ double kepler_sepa_from_freq(const EvoConsts& K, double mtot, double freq) {
    return K.kepler_sepa * pow(mtot, 1.0 / 3.0) / pow(freq, 2.0 / 3.0);
}

// sam_cyutils.pyx:240-253
__device__ __forceinline__ double hard_func_2pwl_gw(const EvoConsts& K, double mtot, double mrat, double sepa,
                                                    double norm, double rchar, double g_in, double g_out) {
    double xx = sepa / rchar;
    double dadt = -norm * pow(1.0 + xx, -g_out + g_in) / pow(xx, g_in - 1.0);
    return dadt + hard_gw(K, mtot, mrat, sepa);
}

// cyutils.pyx:322-358
__device__ __forceinline__ double interp_between(double xnew, double xl, double xr, double yl, double yr) {
    return yl + (yr - yl) * (xnew - xl) / (xr - xl);
}
__device__ __forceinline__ double interp_at_index(int idx, double xnew, const double* xold, const double* yold) {
    return interp_between(xnew, xold[idx], xold[idx + 1], yold[idx], yold[idx + 1]);
}

// sam_cyutils.pyx:64-84: walk an INCREASING table to the interval bracketing val,
// clamped to [0, size-2].  The result only depends on `start` when val equals
// a node exactly, where both neighbouring intervals interpolate to that node.
__device__ __forceinline__ int ww_increasing(int start, int size, double val, const double* edges) {
    int index = start;
    while ((index < size - 2) && (edges[index + 1] < val)) index += 1;
    while ((index > 0) && (edges[index] > val)) index -= 1;
    return index;
}
// sam_cyutils.pyx:87-107 (DECREASING table); clamped to size-2 so that the
// interpolation never reads past the table (the reference bound is size-1,
// reachable only for val below the last node).
__device__ __forceinline__ int ww_decreasing(int start, int size, double val, const double* edges) {
    int index = start;
    while ((index < size - 2) && (edges[index + 1] > val)) index += 1;
    while ((index > 0) && (edges[index - 1] < val)) index -= 1;
    return index;
}

// ---- flat LambdaCDM without radiation (the reference's astropy-based cosmology) ----
struct Cosmo {
    double om0;          // matter density
    double ode0;         // 1 - om0
    double t_hubble;     // 1/H0 [s]
    double d_hubble;     // c/H0 [cm]
    double t_now;        // age at z=0 [s]
    const double* dc_tab;  // (n_tab+1) comoving distance / d_hubble at u = k/n_tab, u = 1/sqrt(1+z)
    const double* dc_der;  // derivative d(dc/d_hubble)/du at the same nodes
    int n_tab;
};

// age(z) = t_H * 2/(3 sqrt(Ode)) * asinh( sqrt(Ode/Om) (1+z)^-1.5 )
__device__ __forceinline__ double cosmo_age(const Cosmo& C, double z) {
    double zp1 = 1.0 + z;
    double x = sqrt(C.ode0 / C.om0) / (zp1 * sqrt(zp1));
    return C.t_hubble * (2.0 / 3.0) / sqrt(C.ode0) * asinh(x);
}
// inverse of cosmo_age
__device__ __forceinline__ double cosmo_tage_to_z(const Cosmo& C, double t) {
    double s = sinh(1.5 * sqrt(C.ode0) * t / C.t_hubble);
    return pow(sqrt(C.ode0 / C.om0) / s, 2.0 / 3.0) - 1.0;
}
// dt/dz magnitude = 1 / ((1+z) H(z))  [s]
__device__ __forceinline__ double cosmo_dtdz(const Cosmo& C, double z) {
    double zp1 = 1.0 + z;
    double E = sqrt(C.om0 * zp1 * zp1 * zp1 + C.ode0);
    return C.t_hubble / (zp1 * E);
}
// comoving distance [cm] from a cubic Hermite table uniform in u = (1+z)^-1/2
__device__ __forceinline__ double cosmo_dcom(const Cosmo& C, double z) {
    double u = 1.0 / sqrt(1.0 + z);
    double x = u * C.n_tab;
    int k = (int)floor(x);
    if (k >= C.n_tab) k = C.n_tab - 1;
    if (k < 0) k = 0;
    double t = x - k;
    double h = 1.0 / C.n_tab;
    double y0 = C.dc_tab[k], y1 = C.dc_tab[k + 1];
    double d0 = C.dc_der[k] * h, d1 = C.dc_der[k + 1] * h;
    double t2 = t * t, t3 = t2 * t;
    double v = (2 * t3 - 3 * t2 + 1) * y0 + (t3 - 2 * t2 + t) * d0 + (-2 * t3 + 3 * t2) * y1 + (t3 - t2) * d1;
    return v * C.d_hubble;
}

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace hd
