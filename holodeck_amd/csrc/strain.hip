// strain.hip -- h_s^2 f/df per bin, <- gravwaves.char_strain_sq_from_bin_edges_redz
// (holodeck/gravwaves.py:694-725).
//
// Per output bin (m,q,z) and frequency f:
//   z_mid  = midpoint of redz_final over the bin's 8 corners, taken axis by axis
//            (m, then q, then z) including -1 sentinels (gravwaves.py:705-708);
//   d_c    = comoving distance(z_mid) for z_mid > 0, else +inf (gravwaves.py:712-714);
//   f_r    = f_c (1 + z_mid)                                   (utils.py:1645-1662);
//   h_s    = K M_c (2 M_c f_r)^(2/3) / d_c                     (utils.py:2260-2285);
//   h2fdf  = h_s^2 (f_c / df).
#include "common.hpp"
#include "../../include/holohip.h"

namespace hd {
Cosmo make_cosmo(double om0, double h, const double* dc_tab, const double* dc_der, int n_tab);

__global__ void __launch_bounds__(256) k_h2fdf(const double* __restrict__ mtc, int M, const double* __restrict__ mrc,
                                               int Q, int Z, const double* __restrict__ fc,
                                               const double* __restrict__ fcdf, int F,
                                               const double* __restrict__ rz, double K, Cosmo C,
                                               double* __restrict__ h2fdf) {
    int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nout = (int64_t)(M - 1) * (Q - 1) * (Z - 1) * F;
    if (o >= nout) return;
    int ff = (int)(o % F);
    int64_t t = o / F;
    int zz = (int)(t % (Z - 1));
    t /= (Z - 1);
    int qq = (int)(t % (Q - 1));
    int mm = (int)(t / (Q - 1));
    auto R = [&](int i, int j, int k) { return rz[((((int64_t)(mm + i) * Q) + (qq + j)) * Z + (zz + k)) * F + ff]; };
    // axis 0 (mass), then axis 1 (ratio), then axis 2 (redshift)
    double a00 = 0.5 * (R(0, 0, 0) + R(1, 0, 0));
    double a01 = 0.5 * (R(0, 0, 1) + R(1, 0, 1));
    double a10 = 0.5 * (R(0, 1, 0) + R(1, 1, 0));
    double a11 = 0.5 * (R(0, 1, 1) + R(1, 1, 1));
    double b0 = 0.5 * (a00 + a10);
    double b1 = 0.5 * (a01 + a11);
    double zmid = 0.5 * (b0 + b1);

    double mt = mtc[mm], mr = mrc[qq];
    double mc = mt * pow(mr, 3.0 / 5.0) / pow(1 + mr, 6.0 / 5.0);
    double dc = (zmid > 0.0) ? cosmo_dcom(C, zmid) : __builtin_inf();
    double fr = fc[ff] * (1.0 + zmid);
    double hs = K * mc * pow(2 * mc * fr, 2.0 / 3.0) / dc;
    h2fdf[o] = (hs * hs) * fcdf[ff];
}

// Per-bin parameters for the strain-weighted averages of the loudest walk
// with parameters (single_sources.py:104-139): [mt, mr, rz] bin centres,
// redz_final (midpointed, -1 where not positive), comoving distance [cm]
// (+inf where redz_final <= 0), final separation [cm] (utils.py:1705-1724,
// astropy G) and angular separation (utils.py:1897-1917).
__global__ void __launch_bounds__(256) k_bin_params(const double* __restrict__ mtc, int M, const double* __restrict__ mrc,
                                                    int Q, const double* __restrict__ rzc, int Z,
                                                    const double* __restrict__ fc, int F, const double* __restrict__ rz,
                                                    Cosmo C, double* __restrict__ out) {
    int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t nout = (int64_t)(M - 1) * (Q - 1) * (Z - 1) * F;
    if (o >= nout) return;
    int ff = (int)(o % F);
    int64_t t = o / F;
    int zz = (int)(t % (Z - 1));
    t /= (Z - 1);
    int qq = (int)(t % (Q - 1));
    int mm = (int)(t / (Q - 1));
    auto R = [&](int i, int j, int k) { return rz[((((int64_t)(mm + i) * Q) + (qq + j)) * Z + (zz + k)) * F + ff]; };
    double a00 = 0.5 * (R(0, 0, 0) + R(1, 0, 0));
    double a01 = 0.5 * (R(0, 0, 1) + R(1, 0, 1));
    double a10 = 0.5 * (R(0, 1, 0) + R(1, 1, 0));
    double a11 = 0.5 * (R(0, 1, 1) + R(1, 1, 1));
    double b0 = 0.5 * (a00 + a10);
    double b1 = 0.5 * (a01 + a11);
    double zmid = 0.5 * (b0 + b1);
    bool sel = zmid > 0.0;
    double redz = sel ? zmid : -1.0;
    double dcom = sel ? cosmo_dcom(C, zmid) : __builtin_inf();
    double frst = fc[ff] * (1.0 + redz);
    const double NWTG = 6.6743e-08;
    double w = 2.0 * M_PI * frst;
    double sepa = pow(NWTG * mtc[mm] / (w * w), 1.0 / 3.0);
    double angs = sepa / (dcom / (1.0 + redz));
    int64_t BF = nout;
    out[0 * BF + o] = mtc[mm];
    out[1 * BF + o] = mrc[qq];
    out[2 * BF + o] = rzc[zz];
    out[3 * BF + o] = redz;
    out[4 * BF + o] = dcom;
    out[5 * BF + o] = sepa;
    out[6 * BF + o] = angs;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_bin_params_redz(const double* mt_cent, int M, const double* mr_cent, int Q, const double* rz_cent,
                                  int Z, const double* fc, int F, const double* redz_final, double om0, double h,
                                  const double* dc_tab, const double* dc_der, int n_tab, double* out, void* stream) {
    if (M < 2 || Q < 2 || Z < 2 || F < 1 || !dc_tab || !dc_der || n_tab < 8 || !out)
        return set_error(HD_ERR_ARG, "hd_bin_params_redz: bad arguments");
    int64_t n = (int64_t)(M - 1) * (Q - 1) * (Z - 1) * F;
    Cosmo C = make_cosmo(om0, h, dc_tab, dc_der, n_tab);
    hipLaunchKernelGGL(k_bin_params, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, mt_cent, M,
                       mr_cent, Q, rz_cent, Z, fc, F, redz_final, C, out);
    return check_launch("k_bin_params");
}

extern "C" int hd_h2fdf_redz(const double* mt_cent, int M, const double* mr_cent, int Q, int Z, const double* fc,
                             const double* fc_over_df, int F, const double* redz_final, double gw_src_const,
                             double om0, double h, const double* dc_tab, const double* dc_der, int n_tab,
                             double* h2fdf, void* stream) {
    if (M < 2 || Q < 2 || Z < 2 || F < 1 || !dc_tab || !dc_der || n_tab < 8)
        return set_error(HD_ERR_ARG, "hd_h2fdf_redz: bad arguments");
    int64_t n = (int64_t)(M - 1) * (Q - 1) * (Z - 1) * F;
    Cosmo C = make_cosmo(om0, h, dc_tab, dc_der, n_tab);
    hipLaunchKernelGGL(k_h2fdf, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, mt_cent, M, mr_cent, Q,
                       Z, fc, fc_over_df, F, redz_final, gw_src_const, C, h2fdf);
    return check_launch("k_h2fdf");
}
