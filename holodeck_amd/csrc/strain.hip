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

// K M_c (2 M_c)^(2/3) of mass bin mm, ratio bin qq: the factor of h_s that depends on the bin
// only, so that h_s = g (f_r)^(2/3) / d_c with one cube root per output
__device__ __forceinline__ double h2_mass_factor(const double* mtc, const double* mrc, int mm, int qq, double K) {
    const double mt = mtc[mm], mr = mrc[qq];
    const double mc = mt * pow(mr, 3.0 / 5.0) / pow(1 + mr, 6.0 / 5.0);
    return K * mc * cbrt((2 * mc) * (2 * mc));
}

// block b: (mass, ratio) bin b / nchunk, chunk b % nchunk of its 256-output chunks of (redshift
// bin, frequency), the chunks of a bin consecutive; one 32-bit division per thread
__global__ void __launch_bounds__(256) k_h2fdf(const double* __restrict__ mtc, int M, const double* __restrict__ mrc,
                                               int Q, int Z, const double* __restrict__ fc,
                                               const double* __restrict__ fcdf, int F,
                                               const double* __restrict__ rz, double K, Cosmo C,
                                               double* __restrict__ h2fdf) {
    const uint32_t per_mq = (uint32_t)(Z - 1) * (uint32_t)F;
    const uint32_t nchunk = (per_mq + blockDim.x - 1) / blockDim.x;
    const int mq = (int)(blockIdx.x / nchunk);   // uniform: the block's chunks of one bin are consecutive
    const int mm = mq / (Q - 1), qq = mq - mm * (Q - 1);
    const uint32_t loc = (blockIdx.x - (uint32_t)mq * nchunk) * blockDim.x + threadIdx.x;
    // the bin's mass factor (two powers and a cube root) once per block, not once per wave; the
    // other waves meanwhile load their corners and evaluate d_c and the cube root, and wait for it
    // only where it is used
    __shared__ double s_g;
    if (threadIdx.x == 0) s_g = h2_mass_factor(mtc, mrc, mm, qq, K);
    const bool live = loc < per_mq;
    const int zz = (int)(loc / (uint32_t)F), ff = (int)(loc - (uint32_t)zz * (uint32_t)F);
    double dc = 1.0, t = 0.0;
    if (live) {
        auto R = [&](int i, int j, int k) { return rz[((((int64_t)(mm + i) * Q) + (qq + j)) * Z + (zz + k)) * F + ff]; };
        // axis 0 (mass), then axis 1 (ratio), then axis 2 (redshift)
        double a00 = 0.5 * (R(0, 0, 0) + R(1, 0, 0));
        double a01 = 0.5 * (R(0, 0, 1) + R(1, 0, 1));
        double a10 = 0.5 * (R(0, 1, 0) + R(1, 1, 0));
        double a11 = 0.5 * (R(0, 1, 1) + R(1, 1, 1));
        double b0 = 0.5 * (a00 + a10);
        double b1 = 0.5 * (a01 + a11);
        double zmid = 0.5 * (b0 + b1);
        dc = (zmid > 0.0) ? cosmo_dcom(C, zmid) : __builtin_inf();
        const double fr = fc[ff] * (1.0 + zmid);
        t = cbrt(fr * fr);
    }
    __syncthreads();
    if (!live) return;
    const double hs = s_g * t / dc;   // K M_c (2 M_c f_r)^(2/3) / d_c
    h2fdf[(int64_t)mq * per_mq + loc] = (hs * hs) * fcdf[ff];
}

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
    const int64_t per_mq = (int64_t)(Z - 1) * F;
    const int64_t nblk = (int64_t)(M - 1) * (Q - 1) * cdiv(per_mq, 256);
    if (per_mq > 0xffffffffll || nblk > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_h2fdf_redz: grid too large");
    Cosmo C = make_cosmo(om0, h, dc_tab, dc_der, n_tab);
    hipLaunchKernelGGL(k_h2fdf, dim3((unsigned)nblk), dim3(256), 0,
                       (hipStream_t)stream, mt_cent, M, mr_cent, Q,
                       Z, fc, fc_over_df, F, redz_final, gw_src_const, C, h2fdf);
    return check_launch("k_h2fdf");
}

// ---------------------------------------------------------------------------
// Idealised, continuous GWB (Phinney 2001, Eq. 5): gravwaves.gwb_ideal
// (gravwaves.py:619-663) as called by Semi_Analytic_Model.gwb_ideal (sam.py:837-870).
//   integrand(i,j,k) = n(i,j,k) [cm^-3] (G Mc)^(5/3) (1+z)^(-1/3), zero where z <= 0,
//   trapezoid over log10 M, q and z, one axis after the other as the reference does; with a
//   per-corner redshift z'(i,j,k) the z widths of a cell come from z' averaged over the cell's
//   four (M, q) corners (the reference's down-averaging of the z coordinates).
//   h_c(f) = sqrt(4 pi / (3 c^2) (pi f)^(-4/3) * integral).
// ---------------------------------------------------------------------------
namespace hd {

constexpr int IDEAL_THREADS = 256;

__device__ __forceinline__ double ideal_integrand(const double* nd, const double* mt, const double* mr,
                                                  const double* rz, bool rz3, int Q, int Z, int i, int j, int k,
                                                  double nwtg, double mpc3) {
    const int64_t c = ((int64_t)i * Q + j) * Z + k;
    const double z = rz3 ? rz[c] : rz[k];
    if (z <= 0.0) return 0.0;
    const double q = mr[j];
    const double mc = mt[i] * pow(q, 3.0 / 5.0) / pow(1.0 + q, 6.0 / 5.0);
    return (nd[c] / mpc3) * pow(nwtg * mc, 5.0 / 3.0) * pow(1.0 + z, -1.0 / 3.0);
}

// per cell: the trapezoid contribution (8-corner mean x dlog10 M x dq x dz)
__device__ __forceinline__ double ideal_cell(const double* nd, const double* lmt, const double* mt, const double* mr,
                                             const double* rz, bool rz3, int Q, int Z, int i, int j, int k,
                                             double nwtg, double mpc3) {
    double s = 0.0;
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
            for (int dk = 0; dk < 2; dk++)
                s += ideal_integrand(nd, mt, mr, rz, rz3, Q, Z, i + di, j + dj, k + dk, nwtg, mpc3);
    double dz;
    if (rz3) {
        auto avg4 = [&](int kk) {
            const double a = 0.5 * (rz[((int64_t)i * Q + j) * Z + kk] + rz[((int64_t)(i + 1) * Q + j) * Z + kk]);
            const double b = 0.5 * (rz[((int64_t)i * Q + j + 1) * Z + kk] + rz[((int64_t)(i + 1) * Q + j + 1) * Z + kk]);
            return 0.5 * (a + b);
        };
        dz = avg4(k + 1) - avg4(k);
    } else {
        dz = rz[k + 1] - rz[k];
    }
    return (s / 8.0) * (lmt[i + 1] - lmt[i]) * (mr[j + 1] - mr[j]) * dz;
}

__global__ void __launch_bounds__(IDEAL_THREADS) k_gwb_ideal_sum(const double* __restrict__ nd,
                                                                 const double* __restrict__ lmt,
                                                                 const double* __restrict__ mt, int M,
                                                                 const double* __restrict__ mr, int Q,
                                                                 const double* __restrict__ rz, int rz3, int Z,
                                                                 double nwtg, double mpc3,
                                                                 double* __restrict__ partial) {
    const int64_t ncell = (int64_t)(M - 1) * (Q - 1) * (Z - 1);
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * IDEAL_THREADS + threadIdx.x; t < ncell;
         t += (int64_t)gridDim.x * IDEAL_THREADS) {
        const int k = (int)(t % (Z - 1));
        const int64_t r = t / (Z - 1);
        const int j = (int)(r % (Q - 1));
        const int i = (int)(r / (Q - 1));
        acc += ideal_cell(nd, lmt, mt, mr, rz, rz3 != 0, Q, Z, i, j, k, nwtg, mpc3);
    }
    __shared__ double red[IDEAL_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < IDEAL_THREADS / 64; w++) t += red[w];
        partial[blockIdx.x] = t;
    }
}

// one workgroup: sum the partials, then h_c per frequency
__global__ void __launch_bounds__(IDEAL_THREADS) k_gwb_ideal_finish(const double* __restrict__ partial, int np,
                                                                    const double* __restrict__ fobs, int F,
                                                                    double cst, double* __restrict__ out) {
    __shared__ double red[IDEAL_THREADS / 64];
    __shared__ double total;
    double acc = 0.0;
    for (int p = threadIdx.x; p < np; p += IDEAL_THREADS) acc += partial[p];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < IDEAL_THREADS / 64; w++) t += red[w];
        total = t;
    }
    __syncthreads();
    for (int f = threadIdx.x; f < F; f += IDEAL_THREADS)
        out[f] = sqrt(cst * pow(M_PI * fobs[f], -4.0 / 3.0) * total);
}

// per (cell, frequency): sqrt(const (pi f)^(-4/3) cell integral)  (sum=False)
__global__ void __launch_bounds__(IDEAL_THREADS) k_gwb_ideal_cells(const double* __restrict__ nd,
                                                                   const double* __restrict__ lmt,
                                                                   const double* __restrict__ mt, int M,
                                                                   const double* __restrict__ mr, int Q,
                                                                   const double* __restrict__ rz, int rz3, int Z,
                                                                   const double* __restrict__ fobs, int F, double cst,
                                                                   double nwtg, double mpc3, double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * IDEAL_THREADS + threadIdx.x;
    const int64_t ncell = (int64_t)(M - 1) * (Q - 1) * (Z - 1);
    if (t >= ncell) return;
    const int k = (int)(t % (Z - 1));
    const int64_t r = t / (Z - 1);
    const int j = (int)(r % (Q - 1));
    const int i = (int)(r / (Q - 1));
    const double v = ideal_cell(nd, lmt, mt, mr, rz, rz3 != 0, Q, Z, i, j, k, nwtg, mpc3);
    for (int f = 0; f < F; f++) out[t * F + f] = sqrt(cst * pow(M_PI * fobs[f], -4.0 / 3.0) * v);
}

}  // namespace hd

extern "C" size_t hd_gwb_ideal_workspace_size(void) { return 1024 * sizeof(double) + 256; }

extern "C" int hd_gwb_ideal(const double* ndens, const double* log10_mtot, const double* mtot, int M, const double* mrat,
                            int Q, const double* redz, int redz_3d, int Z, const double* fobs_gw, int F, double nwtg,
                            double splc, double mpc, int sum, double* out, void* workspace, size_t ws_bytes,
                            void* stream) {
    if (!ndens || !log10_mtot || !mtot || !mrat || !redz || !fobs_gw || !out || M < 2 || Q < 2 || Z < 2 || F < 1)
        return set_error(HD_ERR_ARG, "hd_gwb_ideal: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const double cst = (4.0 * M_PI) / (3 * splc * splc);
    const double mpc3 = mpc * mpc * mpc;
    const int64_t ncell = (int64_t)(M - 1) * (Q - 1) * (Z - 1);
    if (!sum) {
        hipLaunchKernelGGL(k_gwb_ideal_cells, dim3((unsigned)cdiv(ncell, IDEAL_THREADS)), dim3(IDEAL_THREADS), 0, st,
                           ndens, log10_mtot, mtot, M, mrat, Q, redz, redz_3d, Z, fobs_gw, F, cst, nwtg, mpc3, out);
        return check_launch("k_gwb_ideal_cells");
    }
    if (!workspace || ws_bytes < hd_gwb_ideal_workspace_size())
        return set_error(HD_ERR_SIZE, "hd_gwb_ideal: workspace too small");
    const int np = (int)min((int64_t)1024, cdiv(ncell, IDEAL_THREADS));
    double* partial = (double*)workspace;
    hipLaunchKernelGGL(k_gwb_ideal_sum, dim3((unsigned)np), dim3(IDEAL_THREADS), 0, st, ndens, log10_mtot, mtot, M, mrat,
                       Q, redz, redz_3d, Z, nwtg, mpc3, partial);
    int rc = check_launch("k_gwb_ideal_sum");
    if (rc) return rc;
    hipLaunchKernelGGL(k_gwb_ideal_finish, dim3(1), dim3(IDEAL_THREADS), 0, st, partial, np, fobs_gw, F, cst, out);
    return check_launch("k_gwb_ideal_finish");
}
