// cosmo.hip -- flat LambdaCDM quantities in FP64 on the device.
//
// The reference gets these from a cosmopy/astropy Cosmology instance
// (holodeck/__init__.py:81-86): closed-form age and dt/dz, and the comoving
// distance integral.  Here the comoving distance is tabulated once per
// cosmology as a cubic Hermite table in u = (1+z)^(-1/2), where
//     d_c = (c/H0) * integral_u^1  2 du' / sqrt(Om + (1-Om) u'^6)
// has an analytic integrand on [0,1]; each node is integrated with 24-point
// Gauss-Legendre (relative error below 1e-15), and the table interpolates to
// better than 1e-13 relative at n_tab = 4096.
#include "common.hpp"
#include "../../include/holohip.h"

namespace hd {

// 24-point Gauss-Legendre nodes/weights on [-1,1] (positive half).
__constant__ double GL24_X[12] = {
    0.0640568928626056260850430826247450385909, 0.1911188674736163091586398207570696318404,
    0.3150426796961633743867932913198102407864, 0.4337935076260451384870842319133497124524,
    0.5454214713888395356583756172183723700107, 0.6480936519369755692524957869107476266696,
    0.7401241915785543642438281030999784255232, 0.8200019859739029219539498726697452080761,
    0.8864155270044010342131543419821967550873, 0.9382745520027327585236490017087214496548,
    0.9747285559713094981983919930081690617411, 0.9951872199970213601799974097007368118745};
__constant__ double GL24_W[12] = {
    0.1279381953467521569740561652246953718517, 0.1258374563468282961213753825111836887264,
    0.1216704729278033912044631534762624256070, 0.1155056680537256013533444839067835598622,
    0.1074442701159656347825773424466062227946, 0.0976186521041138882698806644642471544279,
    0.0861901615319532759171852029837426671850, 0.0733464814110803057340336152531165181193,
    0.0592985849154367807463677585001085845412, 0.0442774388174198061686027482113382288593,
    0.0285313886289336631813078159518782864491, 0.0123412297999871995468056670700372915759};

__device__ double dc_integrand(double om0, double u) {
    double u2 = u * u, u6 = u2 * u2 * u2;
    return 2.0 / sqrt(om0 + (1.0 - om0) * u6);
}

// integral_u^1 of the integrand (units of c/H0)
__device__ double dc_quad(double om0, double u) {
    double a = u, b = 1.0;
    double c = 0.5 * (b - a), m = 0.5 * (b + a);
    double s = 0.0;
    for (int i = 0; i < 12; i++) {
        s += GL24_W[i] * (dc_integrand(om0, m + c * GL24_X[i]) + dc_integrand(om0, m - c * GL24_X[i]));
    }
    return s * c;
}

__global__ void k_dcom_table(double om0, int n_tab, double* tab, double* der) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n_tab) return;
    double u = (double)k / n_tab;
    tab[k] = dc_quad(om0, u);
    der[k] = -dc_integrand(om0, u);
}

__global__ void k_cosmo_eval(Cosmo C, const double* z, int64_t n, double* age, double* dcom, double* dtdz) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double zz = z[i];
    if (age) age[i] = cosmo_age(C, zz);
    if (dcom) dcom[i] = cosmo_dcom(C, zz);
    if (dtdz) dtdz[i] = cosmo_dtdz(C, zz);
}

}  // namespace hd

using namespace hd;

// Build the Cosmo parameter block (host side).  Hubble time and distance in
// cgs from h, with astropy's Mpc (3.0856775814913673e24 cm) and c.
namespace hd {
Cosmo make_cosmo(double om0, double h, const double* dc_tab, const double* dc_der, int n_tab) {
    const double MPC_CM = 3.0856775814913673e24;
    const double SPLC = 2.99792458e10;
    Cosmo C;
    C.om0 = om0;
    C.ode0 = 1.0 - om0;
    double H0 = h * 100.0 * 1.0e5 / MPC_CM;   // [1/s]
    C.t_hubble = 1.0 / H0;
    C.d_hubble = SPLC / H0;
    C.t_now = C.t_hubble * (2.0 / 3.0) / sqrt(C.ode0) * asinh(sqrt(C.ode0 / C.om0));
    C.dc_tab = dc_tab;
    C.dc_der = dc_der;
    C.n_tab = n_tab;
    return C;
}
}  // namespace hd

extern "C" int hd_cosmo_build_dcom_table(double om0, int n_tab, double* dc_tab, double* dc_der, void* stream) {
    if (!(om0 > 0.0 && om0 <= 1.0) || n_tab < 8 || !dc_tab || !dc_der)
        return set_error(HD_ERR_ARG, "hd_cosmo_build_dcom_table: bad arguments (om0=%g n_tab=%d)", om0, n_tab);
    hipLaunchKernelGGL(k_dcom_table, dim3((unsigned)cdiv(n_tab + 1, 256)), dim3(256), 0, (hipStream_t)stream, om0, n_tab,
                       dc_tab, dc_der);
    return check_launch("k_dcom_table");
}

extern "C" int hd_cosmo_eval(double om0, double h, const double* dc_tab, const double* dc_der, int n_tab, const double* z,
                             int64_t n, double* age, double* dcom, double* dtdz, void* stream) {
    if (n < 0 || !z || (dcom && (!dc_tab || !dc_der || n_tab < 8)))
        return set_error(HD_ERR_ARG, "hd_cosmo_eval: bad arguments");
    if (n == 0) return HD_OK;
    Cosmo C = make_cosmo(om0, h, dc_tab, dc_der, n_tab);
    hipLaunchKernelGGL(k_cosmo_eval, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, C, z, n, age, dcom,
                       dtdz);
    return check_launch("k_cosmo_eval");
}
