// eccen.hip -- eccentric SAM GWB: harmonic sums over a single eccentricity evolution e(a).
//
// Reference: cyutils.pyx:361-597 (sam_calc_gwb_single_eccen, deterministic (F, H) strain per
// frequency and harmonic) and cyutils.pyx:600-851 (the _discrete variant: a Poisson count per
// (grid node, frequency, harmonic, realization)), with g(n, e) of cyutils.pyx:145-199.
//
// The reference walks, per (redshift, total mass), the (frequency, harmonic) pairs in
// ascending observer frequency with an evolution index that only moves forward, skips pairs
// below the first evolution point and stops above the last.  Because the evolution frequencies
// pref[e] * sqrt(M) increase with e, that index equals a fresh search: the first e in
// [0, E-2] with f <= pref[e+1] sqrt(M) (E-2 if none); a pair contributes iff
// pref[0] sqrt(M) <= f <= pref[E-1] sqrt(M).  The same products are compared here, so each
// (mass, redshift, frequency, harmonic) is independent work.
//
// Factorisation used by the deterministic kernel: the term of node (i, j, k) and pair (f, n) is
//   A(i, k, f, n) * P(i, j, k),  A = -a^4 / (C_dadt F(e)) g(n, e) (4 / n^2) f_r^(4/3),
//   P = n_c 4 pi c d_c^2 (1+z) (G_src Mc (2 Mc)^(2/3) / d_c)^2 w / (m1 m2 M),
// where the eccentricity (and so A) does not depend on the mass ratio: the GWB of pair (f, n)
// is sum_{i,k} A(i,k,f,n) * sum_j P(i,j,k).  (Multiplication order differs from the
// reference's; the difference is at the 1e-15 level.)
//
// Bessel functions: J_{n-2}(x) and J_{n-1}(x) (x = n e) by Miller's backward recurrence
// (J. C. P. Miller, British Association Mathematical Tables X, 1952; the start order and
// normalisation J_0 + 2 sum J_2k = 1 as in Numerical Recipes' bessj), rescaled against
// overflow; then J_n .. J_{n+2} by the upward recurrence exactly as the reference does, with
// its cancellation at small e kept.
#include "common.hpp"
#include "../../include/holohip.h"

namespace hd {

constexpr double ECC_ZERO_LIMIT = 1.0e-12;   // cyutils.pyx:39
constexpr int ECC_THREADS = 256;
constexpr int ECC_MAX_ORDER = 100000;        // Bessel orders and harmonics accepted (loop length ~ order)
constexpr double ECC_MAX_ARG = 1.0e5;        // Bessel arguments accepted

struct EccConsts {
    double gw_dadt_sep;    // -64 G^3 / 5 / c^5 (cyutils.pyx:47)
    double gw_src;         // 8 G^(5/3) pi^(2/3) / sqrt(10) / c^4 (cyutils.pyx:48)
    double four_pi_c_mpc;  // 4 pi c / Mpc
    double kep_sa_term;    // G / (2 pi)^2
    double mpc;            // Mpc [cm]
};

static EccConsts ecc_consts() {
    EccConsts K;
    K.gw_dadt_sep = -64.0 * pow(MY_NWTG, 3) / 5.0 / pow(MY_SPLC, 5);
    K.gw_src = 8.0 * pow(MY_NWTG, 5.0 / 3.0) * pow(M_PI, 2.0 / 3.0) / sqrt(10.0) / pow(MY_SPLC, 4.0);
    K.four_pi_c_mpc = 4 * M_PI * (MY_SPLC / MY_MPC);
    K.kep_sa_term = MY_NWTG / pow(2.0 * M_PI, 2);
    K.mpc = MY_MPC;
    return K;
}

// J_lo(x), J_{lo+1}(x) for -1 <= lo <= ECC_MAX_ORDER (J_{-1} = -J_1), 0 < x <= ECC_MAX_ARG; NaN
// outside that range (the backward recurrence runs ~ max(lo, x) steps, so the range bounds it)
__device__ __noinline__ void bessel_pair(int lo, double x, double& j_lo, double& j_hi) {
    if (lo < -1 || lo > ECC_MAX_ORDER || !(x > 0.0) || !(x <= ECC_MAX_ARG)) {
        j_lo = j_hi = __builtin_nan("");
        return;
    }
    const int a = lo < 0 ? 0 : lo;                         // orders a, a+1 are computed
    const double big = fmax((double)(a + 2), x);
    const int m = 2 * ((int)(big + sqrt(160.0 * big) + 10.0) / 2);
    const double tox = 2.0 / x;
    double jp = 0.0, j = 1e-30, s = 0.0, c_a = 0.0, c_a1 = 0.0;
    for (int k = m; k > 0; --k) {
        const double jm = ((double)k * tox) * j - jp;    // J_{k-1}
        jp = j;
        j = jm;
        if (k - 1 == a + 1) c_a1 = j;
        if (k - 1 == a) c_a = j;
        if (((k - 1) & 1) == 0 && k > 1) s += 2.0 * j;
        if (fabs(j) > 1e250) {
            j *= 1e-250;
            jp *= 1e-250;
            s *= 1e-250;
            c_a *= 1e-250;
            c_a1 *= 1e-250;
        }
    }
    const double norm = j + s;
    if (lo < 0) {                                          // (J_{-1}, J_0) = (-J_1, J_0)
        j_lo = -c_a1 / norm;
        j_hi = c_a / norm;
    } else {
        j_lo = c_a / norm;
        j_hi = c_a1 / norm;
    }
}

// g(n, e), cyutils.pyx:145-199 (no fused multiply-adds: the upward recurrence cancels at small e
// and the reference's compiled Cython rounds each operation)
__device__ double gw_freq_dist(int nn, double ee) {
#pragma clang fp contract(off)
    if (ee < ECC_ZERO_LIMIT) return nn == 2 ? 1.0 : 0.0;
    const double ne = nn * ee;
    const double n2 = (double)nn * nn;
    double jn_m2, jn_m1;
    bessel_pair(nn - 2, ne, jn_m2, jn_m1);
    const double jn = (2 * (nn - 1) / ne) * jn_m1 - jn_m2;
    const double jn_p1 = (2 * nn / ne) * jn - jn_m1;
    const double jn_p2 = (2 * (nn + 1) / ne) * jn_p1 - jn;
    double aa = jn_m2 - 2.0 * ee * jn_m1 + (2.0 / nn) * jn + 2 * ee * jn_p1 - jn_p2;
    aa = aa * aa;
    double bb = jn_m2 - 2 * ee * jn + jn_p2;
    bb = (1 - ee * ee) * bb * bb;
    const double cc = (4.0 / (3.0 * n2)) * jn * jn;
    return (n2 * n2 / 32) * (aa + bb + cc);
}

// F(e), cyutils.pyx:79-99
__device__ __forceinline__ double gw_ecc_func(double e) {
    const double e2 = e * e;
    return (1.0 + (73.0 / 24.0) * e2 + (37.0 / 96.0) * e2 * e2) / pow(1.0 - e2, 7.0 / 2.0);
}

// cyutils.pyx:102-142: (inverse weight, width) of grid point `index`
__device__ __forceinline__ void trapz_weight(int index, int size, const double* grid, double& w0, double& w1) {
    if (index == 0) {
        w0 = 2.0;
        w1 = grid[1] - grid[0];
    } else if (index == size - 1) {
        w0 = 2.0;
        w1 = grid[index] - grid[index - 1];
    } else {
        w0 = 1.0;
        w1 = 0.5 * (grid[index + 1] - grid[index - 1]);
    }
}

struct EccPair {
    bool ok;
    double gwfr, sa4, ecc;
};

// evolution bracket and interpolated eccentricity for rest-frame frequency gwfr at total mass
// mt (cyutils.pyx:531-556); pref[e] = (sqrt(G)/2pi) / sepa[e]^1.5, increasing
__device__ EccPair ecc_at(const EccConsts& K, const double* pref, const double* eccen, int E, double mt,
                          double mt_sqrt, double gwfr) {
    EccPair r;
    r.gwfr = gwfr;
    int lo = 0, hi = E - 2;                         // first e with gwfr <= pref[e+1] sqrt(mt), else E-2
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (gwfr > pref[mid + 1] * mt_sqrt) lo = mid + 1;
        else hi = mid;
    }
    const double f_lo = pref[lo] * mt_sqrt;
    const double f_hi = pref[lo + 1] * mt_sqrt;
    r.ok = !(gwfr < f_lo) && !(gwfr > f_hi);
    double ecc = (eccen[lo + 1] - eccen[lo]) / (f_hi - f_lo);
    r.ecc = eccen[lo] + (gwfr - f_lo) * ecc;
    const double sa = pow(K.kep_sa_term * mt / pow(gwfr, 2), 1.0 / 3.0);
    r.sa4 = pow(sa, 4);
    return r;
}

// per (i, k): sum over mass ratios of P(i, j, k)
__global__ void k_eccen_psum(const double* __restrict__ ndens, const double* __restrict__ mtot_log10, int M,
                             const double* __restrict__ mrat, int Q, const double* __restrict__ redz,
                             const double* __restrict__ dcom, int Z, EccConsts K, double* __restrict__ psum) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= M * Z) return;
    const int ii = p / Z, kk = p % Z;
    double i0, i1, k0, k1;
    trapz_weight(ii, M, mtot_log10, i0, i1);
    trapz_weight(kk, Z, redz, k0, k1);
    const double zterm = 1.0 + redz[kk];
    const double dc_mpc = dcom[kk];
    const double dc_cm = dc_mpc * K.mpc;
    const double dc_term = K.four_pi_c_mpc * pow(dc_mpc, 2);
    const double mt = pow(10.0, mtot_log10[ii]);
    const double weight_ik = i1 * k1 / (i0 * k0);
    double acc = 0.0;
    for (int jj = 0; jj < Q; ++jj) {
        double j0, j1;
        trapz_weight(jj, Q, mrat, j0, j1);
        const double q = mrat[jj];
        const double m1 = mt / (1.0 + q);
        const double m2 = mt - m1;
        const double mchirp = mt * pow(q, 3.0 / 5.0) / pow(1 + q, 6.0 / 5.0);
        double hp = ndens[((int64_t)ii * Q + jj) * Z + kk] * dc_term * zterm;
        hp *= pow(K.gw_src * mchirp * pow(2.0 * mchirp, 2.0 / 3.0) / dc_cm, 2);
        acc += hp * (weight_ik * (j1 / j0)) / (m1 * m2 * mt);
    }
    psum[p] = acc;
}

// one workgroup per (frequency, harmonic): sum over (i, k) of A * psum
__global__ void __launch_bounds__(ECC_THREADS) k_eccen_gwb(const double* __restrict__ mtot_log10, int M,
                                                           const double* __restrict__ redz, int Z,
                                                           const double* __restrict__ gwfobs, int H,
                                                           const double* __restrict__ pref,
                                                           const double* __restrict__ eccen, int E, EccConsts K,
                                                           const double* __restrict__ psum,
                                                           double* __restrict__ gwb) {
    const int fh = blockIdx.x;
    const int aa = fh / H, nh = fh % H + 1;
    double acc = 0.0;
    for (int p = threadIdx.x; p < M * Z; p += ECC_THREADS) {
        const int ii = p / Z, kk = p % Z;
        const double ps = psum[p];
        if (ps == 0.0) continue;
        const double mt = pow(10.0, mtot_log10[ii]);
        const double gwfr = gwfobs[aa] * (1.0 + redz[kk]) / nh;
        const EccPair e = ecc_at(K, pref, eccen, E, mt, sqrt(mt), gwfr);
        if (!e.ok) continue;
        const double gne = gw_freq_dist(nh, e.ecc);
        if (gne == 0.0) continue;
        const double A = -e.sa4 / (K.gw_dadt_sep * gw_ecc_func(e.ecc)) * gne * (4.0 / (nh * nh)) *
                         pow(gwfr, 4.0 / 3.0);
        acc += A * ps;
    }
    __shared__ double red[ECC_THREADS / 64];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < ECC_THREADS / 64; ++w) t += red[w];
        gwb[fh] = t;
    }
}

// discrete variant: Poisson mean and strain factor per (node, column); columns are
// (frequency, harmonic h0 .. h0+Hc-1) pairs, c = aa * Hc + (nh - 1 - h0)
__global__ void __launch_bounds__(ECC_THREADS) k_eccen_number(
    const double* __restrict__ ndens, const double* __restrict__ mtot_log10, int M, const double* __restrict__ mrat,
    int Q, const double* __restrict__ redz, const double* __restrict__ dcom, int Z, const double* __restrict__ gwfobs,
    int F, int h0, int Hc, const double* __restrict__ pref, const double* __restrict__ eccen, int E, EccConsts K,
    double* __restrict__ lam, double* __restrict__ hval) {
    const int C = F * Hc;
    const int c = blockIdx.x * ECC_THREADS + threadIdx.x;
    const int p = blockIdx.y;
    if (c >= C) return;
    const int ii = p / Z, kk = p % Z;
    const int aa = c / Hc, nh = h0 + c % Hc + 1;
    double i0, i1, k0, k1;
    trapz_weight(ii, M, mtot_log10, i0, i1);
    trapz_weight(kk, Z, redz, k0, k1);
    const double zterm = 1.0 + redz[kk];
    const double dc_mpc = dcom[kk];
    const double dc_cm = dc_mpc * K.mpc;
    const double dc_term = K.four_pi_c_mpc * pow(dc_mpc, 2);
    const double mt = pow(10.0, mtot_log10[ii]);
    const double gwfr = gwfobs[aa] * zterm / nh;
    const EccPair e = ecc_at(K, pref, eccen, E, mt, sqrt(mt), gwfr);
    double gne = 0.0, fe = 1.0, f43 = 0.0;
    if (e.ok) {
        gne = gw_freq_dist(nh, e.ecc);
        fe = gw_ecc_func(e.ecc);
        f43 = pow(gwfr, 4.0 / 3.0);
    }
    for (int jj = 0; jj < Q; ++jj) {
        const int64_t row = ((int64_t)ii * Q + jj) * Z + kk;
        double l = 0.0, h = 0.0;
        if (gne != 0.0) {
            double j0, j1;
            trapz_weight(jj, Q, mrat, j0, j1);
            const double q = mrat[jj];
            const double m1 = mt / (1.0 + q);
            const double m2 = mt - m1;
            const double mchirp = mt * pow(q, 3.0 / 5.0) / pow(1 + q, 6.0 / 5.0);
            const double volume = i1 * k1 * j1;
            const double weight = i0 * k0 * j0;
            const double tau = -e.sa4 / (K.gw_dadt_sep * fe * m1 * m2 * mt);
            const double hp = pow(K.gw_src * mchirp * pow(2.0 * mchirp, 2.0 / 3.0) / dc_cm, 2);
            h = hp * gne * (4.0 / (nh * nh)) * f43 / weight;
            l = (h != 0.0) ? ndens[row] * dc_term * zterm * tau * volume : 0.0;
        }
        lam[row * C + c] = l;
        hval[row * C + c] = h;
    }
}

__global__ void k_eccen_pref(const double* __restrict__ sepa, int E, double* __restrict__ pref) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < E) pref[e] = ((1.0 / (2.0 * M_PI)) * sqrt(MY_NWTG)) / pow(sepa[e], 1.5);
}

__global__ void k_bessel_pair(const int* __restrict__ lo, const double* __restrict__ x, int64_t n,
                              double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double a, b;
    bessel_pair(lo[t], x[t], a, b);
    out[2 * t] = a;
    out[2 * t + 1] = b;
}

__global__ void k_gw_freq_dist(const int* __restrict__ nn, const double* __restrict__ ee, int64_t n,
                               double* __restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = gw_freq_dist(nn[t], ee[t]);
}

static bool ecc_args_ok(const double* ndens, const double* mtot_log10, int M, const double* mrat, int Q,
                        const double* redz, const double* dcom, int Z, const double* gwfobs, int F,
                        const double* sepa, const double* eccen, int E, int nharms) {
    return ndens && mtot_log10 && mrat && redz && dcom && gwfobs && sepa && eccen && M >= 2 && Q >= 2 && Z >= 2 &&
           F >= 1 && E >= 2 && nharms >= 1;
}

}  // namespace hd

using namespace hd;

extern "C" size_t hd_eccen_workspace_size(int M, int Z, int E) {
    return (size_t)(M * (int64_t)Z + E) * sizeof(double) + 256;
}

extern "C" int hd_gwb_eccen(const double* ndens, const double* mtot_log10, int M, const double* mrat, int Q,
                            const double* redz, const double* dcom, int Z, const double* gwfobs, int F,
                            const double* sepa_evo, const double* eccen_evo, int E, int nharms, double* gwb,
                            void* workspace, size_t ws_bytes, void* stream) {
    if (!ecc_args_ok(ndens, mtot_log10, M, mrat, Q, redz, dcom, Z, gwfobs, F, sepa_evo, eccen_evo, E, nharms) || !gwb)
        return set_error(HD_ERR_ARG, "hd_gwb_eccen: bad arguments");
    if (!workspace || ws_bytes < hd_eccen_workspace_size(M, Z, E))
        return set_error(HD_ERR_SIZE, "hd_gwb_eccen: workspace too small");
    if ((int64_t)F * nharms > 0x7fffffff || nharms > ECC_MAX_ORDER)
        return set_error(HD_ERR_SIZE, "hd_gwb_eccen: too many (f, harmonic) pairs (nharms <= %d)", ECC_MAX_ORDER);
    hipStream_t st = (hipStream_t)stream;
    double* psum = (double*)workspace;
    double* pref = psum + (int64_t)M * Z;
    EccConsts K = ecc_consts();
    hipLaunchKernelGGL(k_eccen_pref, dim3((unsigned)cdiv(E, 256)), dim3(256), 0, st, sepa_evo, E, pref);
    hipLaunchKernelGGL(k_eccen_psum, dim3((unsigned)cdiv((int64_t)M * Z, 256)), dim3(256), 0, st, ndens, mtot_log10, M,
                       mrat, Q, redz, dcom, Z, K, psum);
    KTimer timer("k_eccen_gwb", st);
    hipLaunchKernelGGL(k_eccen_gwb, dim3((unsigned)(F * nharms)), dim3(ECC_THREADS), 0, st, mtot_log10, M, redz, Z,
                       gwfobs, nharms, pref, eccen_evo, E, K, psum, gwb);
    return check_launch("k_eccen_gwb");
}

extern "C" int hd_eccen_number(const double* ndens, const double* mtot_log10, int M, const double* mrat, int Q,
                               const double* redz, const double* dcom, int Z, const double* gwfobs, int F,
                               const double* sepa_evo, const double* eccen_evo, int E, int h0, int hc, double* number,
                               double* hval, void* workspace, size_t ws_bytes, void* stream) {
    if (!ecc_args_ok(ndens, mtot_log10, M, mrat, Q, redz, dcom, Z, gwfobs, F, sepa_evo, eccen_evo, E, 1) || !number ||
        !hval || h0 < 0 || hc < 1)
        return set_error(HD_ERR_ARG, "hd_eccen_number: bad arguments");
    if (!workspace || ws_bytes < hd_eccen_workspace_size(M, Z, E))
        return set_error(HD_ERR_SIZE, "hd_eccen_number: workspace too small");
    if ((int64_t)M * Z > 65535) return set_error(HD_ERR_SIZE, "hd_eccen_number: M*Z=%d above 65535", M * Z);
    if ((int64_t)h0 + hc > ECC_MAX_ORDER || (int64_t)F * hc > 0x7fffffff)
        return set_error(HD_ERR_SIZE, "hd_eccen_number: harmonics above %d", ECC_MAX_ORDER);
    hipStream_t st = (hipStream_t)stream;
    double* pref = (double*)workspace + (int64_t)M * Z;
    EccConsts K = ecc_consts();
    hipLaunchKernelGGL(k_eccen_pref, dim3((unsigned)cdiv(E, 256)), dim3(256), 0, st, sepa_evo, E, pref);
    KTimer timer("k_eccen_number", st);
    hipLaunchKernelGGL(k_eccen_number, dim3((unsigned)cdiv((int64_t)F * hc, ECC_THREADS), (unsigned)(M * Z)),
                       dim3(ECC_THREADS), 0, st, ndens, mtot_log10, M, mrat, Q, redz, dcom, Z, gwfobs, F, h0, hc, pref,
                       eccen_evo, E, K, number, hval);
    return check_launch("k_eccen_number");
}

extern "C" int hd_bessel_jpair(const int* lo, const double* x, int64_t n, double* out, void* stream) {
    if (!lo || !x || !out || n < 0) return set_error(HD_ERR_ARG, "hd_bessel_jpair: bad arguments");
    if (n == 0) return HD_OK;
    hipLaunchKernelGGL(k_bessel_pair, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, lo, x, n, out);
    return check_launch("k_bessel_pair");
}

extern "C" int hd_gw_freq_dist(const int* nn, const double* ee, int64_t n, double* out, void* stream) {
    if (!nn || !ee || !out || n < 0) return set_error(HD_ERR_ARG, "hd_gw_freq_dist: bad arguments");
    if (n == 0) return HD_OK;
    hipLaunchKernelGGL(k_gw_freq_dist, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, nn, ee, n,
                       out);
    return check_launch("k_gw_freq_dist");
}
