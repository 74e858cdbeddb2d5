// density.hip -- static binary density per (M,Q,Z) grid corner, FP64.
//
// <- Semi_Analytic_Model.static_binary_density (holodeck/sams/sam.py:280-398)
//    mass_stellar (sam.py:250-278) via MMBulge_Standard / BF_Constant
//    (host_relations.py:166-195, 624-771, 1102-1173), GSMF_Schechter
//    (components.py:110-172), GPF_Power_Law (components.py:520-583),
//    GMT_Power_Law + zprime (components.py:620-675, utils.py:1772-1808),
//    GMR_Illustris (components.py:380-482).
// The M-Mbulge scatter step runs afterwards (scatter.hip), then
// hd_zero_stalled applies `dens[zprime < 0] = 0` (sam.py:392-394).
#include "common.hpp"
#include "../../include/holohip.h"
#include "density_par.h"

namespace hd {
Cosmo make_cosmo(double om0, double h, const double* dc_tab, const double* dc_der, int n_tab);

struct DensPar {
    double p[HD_DENS_NPAR];
};

__device__ __forceinline__ double mstar_from_mbh(const double* P, double mbh) {
    // _log10_relation_reverse (host_relations.py:1137-1173), then BF_Constant
    double xx = log10(mbh / P[DP_MMB_MAMP]);
    xx = (1.0 / P[DP_MMB_PLAW]) * xx;
    double mbulge = P[DP_MMB_MREF] * pow(10.0, xx);
    return mbulge / P[DP_MMB_BFRAC];
}

__global__ void __launch_bounds__(256) k_static_density(const double* __restrict__ mtot, int M,
                                                        const double* __restrict__ mrat, int Q,
                                                        const double* __restrict__ redz, int Z, DensPar DP, Cosmo C,
                                                        double* __restrict__ dens, double* __restrict__ zprime,
                                                        double* __restrict__ gmt_time) {
    int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (int64_t)M * Q * Z) return;
    const double* P = DP.p;
    int kk = (int)(cell % Z);
    int jj = (int)((cell / Z) % Q);
    int ii = (int)(cell / ((int64_t)Q * Z));
    double mt = mtot[ii], mr = mrat[jj], z = redz[kk];

    // mass_stellar: m1 = M/(1+q), m2 = M - m1 (utils.py:1620-1642)
    double m1 = mt / (1.0 + mr);
    double m2 = mt - m1;
    double mstar_pri = mstar_from_mbh(P, m1);
    double mstar_sec = mstar_from_mbh(P, m2);
    double mstar_rat = mstar_sec / mstar_pri;
    double mstar_tot = mstar_pri + mstar_sec;

    // galaxy merger time and delayed redshift
    double tau = 0.0, zp = z;
    const bool has_gmt = P[DP_HAS_GMT] != 0.0;
    if (has_gmt) {
        tau = P[DP_GMT_NORM] * pow(mstar_pri / P[DP_GMT_MREF], P[DP_GMT_MALPHA]) * pow(1.0 + z, P[DP_GMT_ZBETA]) *
              pow(mstar_rat, P[DP_GMT_QGAMMA]);
        double new_age = cosmo_age(C, z) + tau;
        zp = (new_age < C.t_now) ? cosmo_tage_to_z(C, new_age) : -1.0;
    }

    double rate;
    if (P[DP_USE_GMR] == 0.0) {
        double rv = P[DP_GPF_NORM] * pow(mstar_pri / P[DP_GPF_MREF], P[DP_GPF_MALPHA]) * pow(1.0 + z, P[DP_GPF_ZBETA]) *
                    pow(mstar_rat, P[DP_GPF_QGAMMA]);
        rv = fmin(rv, P[DP_GPF_MAXFRAC]);
        rate = rv / tau;
    } else {
        double zp1 = 1.0 + z;
        double norm = P[DP_GMR_NORM0] * pow(zp1, P[DP_GMR_NORMZ]);
        double malpha = P[DP_GMR_MALPHA0] * pow(zp1, P[DP_GMR_MALPHAZ]);
        double mdelta = P[DP_GMR_MDELTA0] * pow(zp1, P[DP_GMR_MDELTAZ]);
        double qgamma = P[DP_GMR_QGAMMA0] * pow(zp1, P[DP_GMR_QGAMMAZ]);
        qgamma = qgamma + P[DP_GMR_QGAMMAM] * log10(mstar_tot / P[DP_GMR_MREF]);
        double xx = mstar_tot / P[DP_GMR_MREF];
        double mtp = pow(xx, malpha);
        double yy = mstar_tot / P[DP_GMR_MREF_DELTA];
        double mp1t = pow(1.0 + yy, mdelta);
        double qt = pow(mstar_rat, qgamma);
        rate = norm * mtp * mp1t * qt;
    }

    // GSMF_Schechter (components.py:132-172); mchar adds mcharz*z to grams as the reference does
    double phi = pow(10.0, P[DP_GSMF_PHI0] + P[DP_GSMF_PHIZ] * z);
    double mchar = P[DP_GSMF_MCHAR0] + P[DP_GSMF_MCHARZ] * z;
    double alpha = P[DP_GSMF_ALPHA0] + P[DP_GSMF_ALPHAZ] * z;
    double xg = mstar_pri / mchar;
    double gsmf = log(10.0) * phi * pow(xg, 1.0 + alpha) * exp(-xg);

    double d = gsmf * rate * cosmo_dtdz(C, z);

    // Jacobian dMstar/dMbh and dq (sam.py:355-365, host_relations.py:483-512, 720-743)
    double plaw = P[DP_MMB_PLAW];
    double dqbh_dqgal = plaw * pow(mstar_rat, plaw - 1.0);
    double mbulge = mstar_pri * P[DP_MMB_BFRAC];
    double mbh = P[DP_MMB_MAMP] * pow(10.0, log10(mbulge / P[DP_MMB_MREF]) * plaw);
    double dmbulge_dmbh = mbulge / (plaw * mbh);
    double dmstar_dmbh_pri = (1.0 / P[DP_MMB_BFRAC]) * dmbulge_dmbh;
    double qterm = (1.0 + mstar_rat) / (1.0 + mr);
    double dmstar_dmbh = dmstar_dmbh_pri * qterm;
    d *= (mt / mstar_tot) * (dmstar_dmbh / dqbh_dqgal);

    dens[cell] = d;
    zprime[cell] = zp;
    gmt_time[cell] = tau;
}

__global__ void k_zero_stalled(double* dens, const double* zprime, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (zprime[i] < 0.0) dens[i] = 0.0;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_static_density(const double* mtot, int M, const double* mrat, int Q, const double* redz, int Z,
                                 const double* par, const double* dc_tab, const double* dc_der, int n_tab,
                                 double* dens, double* zprime, double* gmt_time, void* stream) {
    if (M < 1 || Q < 1 || Z < 1 || !par || !dens || !zprime || !gmt_time)
        return set_error(HD_ERR_ARG, "hd_static_density: bad arguments");
    DensPar DP;
    for (int i = 0; i < HD_DENS_NPAR; i++) DP.p[i] = par[i];   // host array
    if (!(DP.p[DP_COSMO_OM0] > 0.0 && DP.p[DP_COSMO_OM0] < 1.0 && DP.p[DP_COSMO_H] > 0.0))
        return set_error(HD_ERR_ARG, "hd_static_density: bad cosmology parameters");
    Cosmo C = make_cosmo(DP.p[DP_COSMO_OM0], DP.p[DP_COSMO_H], dc_tab, dc_der, n_tab);
    int64_t n = (int64_t)M * Q * Z;
    hipLaunchKernelGGL(k_static_density, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, mtot, M, mrat,
                       Q, redz, Z, DP, C, dens, zprime, gmt_time);
    return check_launch("k_static_density");
}

namespace hd {
// mass_stellar (sam.py:250-278): per (M, q) node, broadcast over the Z redshifts as the reference
__global__ void __launch_bounds__(256) k_mass_stellar(const double* __restrict__ mtot, int M,
                                                      const double* __restrict__ mrat, int Q, int Z, DensPar DP,
                                                      double* __restrict__ mpri, double* __restrict__ mrt,
                                                      double* __restrict__ mtt) {
    const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (int64_t)M * Q * Z) return;
    const int jj = (int)((cell / Z) % Q);
    const int ii = (int)(cell / ((int64_t)Q * Z));
    const double mt = mtot[ii], mr = mrat[jj];
    const double m1 = mt / (1.0 + mr);
    const double m2 = mt - m1;
    const double sp = mstar_from_mbh(DP.p, m1);
    const double ss = mstar_from_mbh(DP.p, m2);
    mpri[cell] = sp;
    mrt[cell] = ss / sp;
    mtt[cell] = sp + ss;
}
}  // namespace hd

extern "C" int hd_mass_stellar(const double* mtot, int M, const double* mrat, int Q, int Z, const double* par,
                               double* mstar_pri, double* mstar_rat, double* mstar_tot, void* stream) {
    if (M < 1 || Q < 1 || Z < 1 || !mtot || !mrat || !par || !mstar_pri || !mstar_rat || !mstar_tot)
        return set_error(HD_ERR_ARG, "hd_mass_stellar: bad arguments");
    DensPar DP;
    for (int i = 0; i < HD_DENS_NPAR; i++) DP.p[i] = par[i];   // host array
    const int64_t n = (int64_t)M * Q * Z;
    hipLaunchKernelGGL(k_mass_stellar, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, mtot, M, mrat,
                       Q, Z, DP, mstar_pri, mstar_rat, mstar_tot);
    return check_launch("k_mass_stellar");
}

extern "C" int hd_zero_stalled(double* dens, const double* zprime, int64_t n, void* stream) {
    if (n < 0 || (n > 0 && (!dens || !zprime))) return set_error(HD_ERR_ARG, "hd_zero_stalled: bad arguments");
    if (n == 0) return HD_OK;
    hipLaunchKernelGGL(k_zero_stalled, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, dens, zprime, n);
    return check_launch("k_zero_stalled");
}
