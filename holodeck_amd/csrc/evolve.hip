// evolve.hip -- binary evolution kernels of the SAM hot path (FP64).
//
//   hd_find_2pwl_norm      <- sam_cyutils.find_2pwl_hardening_norm   (sam_cyutils.pyx:289-398)
//   hd_hard_func_2pwl_gw   <- sam_cyutils.hard_func_2pwl_gw          (sam_cyutils.pyx:256-286)
//   hd_dbn_gw              <- _dynamic_binary_number_at_fobs_gw      (sam_cyutils.pyx:784-899)
//   hd_dbn_2pwl            <- _dynamic_binary_number_at_fobs_2pwl    (sam_cyutils.pyx:506-781)
//   hd_integrate_3dx1d     <- integrate_differential_number_3dx1d    (sam_cyutils.pyx:115-216)
//
// All arithmetic is FP64: the z<->t inversion through the age table and the
// broken-power-law lifetime integral lose stability in FP32.
#include "common.hpp"
#include "../../include/holohip.h"
#include <math.h>

namespace hd {

// ---------------------------------------------------------------------------
// Brent root of the 2PL lifetime, one wavefront per (M,Q) element.
//
// The 300-step trapezoid lifetime (sam_cyutils.pyx:356-398) is spread over the
// 64 lanes in contiguous chunks; each lane reproduces the reference's
// sequential `sepa_log10 -= dx` chain for its chunk, sums its dt's in order and
// the chunk sums are combined with a butterfly.  The Brent iteration itself
// (scalar control flow, identical on all lanes) restates the published
// bracketing algorithm used by scipy's brentq, which the reference calls with
// xtol=1e-3, rtol=1e-5, maxiter=100 on [-20, 20] (sam_cyutils.pyx:332-349).
// ---------------------------------------------------------------------------
// x^e for x > 0: from products and at most one square root when 2e is an integer of magnitude
// up to 32 (the hardening model's exponents, -3.5 and -2 by default; a few ulp), pow otherwise.
// e is the same for every thread, so the branch is uniform.
__device__ __forceinline__ double pow_halfint(double x, double e) {
    const double a = fabs(e);
    if (!(a <= 16.0) || 2.0 * a != floor(2.0 * a)) return pow(x, e);
    int n = (int)a;
    double r = (a - (double)n != 0.0) ? sqrt(x) : 1.0;
    double b = x;
    while (n > 0) {
        if (n & 1) r *= b;
        b *= b;
        n >>= 1;
    }
    return (e < 0.0) ? 1.0 / r : r;
}

struct LifeArgs {
    double mt, mr, target, rchar, g_in, g_out;
    double x_lane0;   // sepa_log10 before this lane's first step
    double dx;
    int s_begin, s_end;  // this lane's steps [s_begin, s_end), 0-based
};

__device__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ double lifetime_wave(const EvoConsts& K, const LifeArgs& a, double norm_log10) {
    double norm = pow(10.0, norm_log10);
    double part = 0.0;
    if (a.s_begin < a.s_end) {
        double x = a.x_lane0;
        double sepa_left = pow(10.0, x);
        double dadt_left = hard_func_2pwl_gw(K, a.mt, a.mr, sepa_left, norm, a.rchar, a.g_in, a.g_out);
        for (int s = a.s_begin; s < a.s_end; s++) {
            x -= a.dx;
            double sepa_right = pow(10.0, x);
            double dadt_right = hard_func_2pwl_gw(K, a.mt, a.mr, sepa_right, norm, a.rchar, a.g_in, a.g_out);
            double dt = 2.0 * (sepa_right - sepa_left) / (dadt_left + dadt_right);
            part += dt;
            sepa_left = sepa_right;
            dadt_left = dadt_right;
        }
    }
    return wave_sum(part) - a.target;
}

// The factors of the hardening rate that do not depend on the normalisation, at this lane's
// step boundaries, computed once per (M, Q) instead of at every root-finder iteration: the
// separation, (1 + x)^(g_in - g_out), x^(g_in - 1) with x = a / r_char, the GW rate, and the
// step's 2 (a_right - a_left).  The rate is then ((-norm) * p1) / p2 + gw, the same operations in
// the same order as hard_func_2pwl_gw, so every lifetime is bit for bit the one computed from
// scratch.  PTS - 1 steps per lane at most (nsteps <= 64 (PTS - 1)); the arrays are sized to the
// launch (6 points for up to 320 steps) so that they stay in few registers.
template <int PTS>
struct LifePre {
    double p1[PTS], p2[PTS], gw[PTS], num[PTS];
};

template <int PTS>
__device__ __forceinline__ void lifetime_pre(const EvoConsts& K, const LifeArgs& a, LifePre<PTS>& P) {
    const int npts = (a.s_begin < a.s_end) ? a.s_end - a.s_begin + 1 : 0;
    double x = a.x_lane0, sepa_prev = 0.0;
#pragma unroll
    for (int i = 0; i < PTS; i++) {
        if (i > 0) x -= a.dx;
        const double sepa = pow(10.0, x);
        const double xx = sepa / a.rchar;
        const bool on = i < npts;
        P.p1[i] = on ? pow(1.0 + xx, -a.g_out + a.g_in) : 1.0;
        P.p2[i] = on ? pow(xx, a.g_in - 1.0) : 1.0;
        P.gw[i] = on ? hard_gw(K, a.mt, a.mr, sepa) : 0.0;
        P.num[i] = (on && i > 0) ? 2.0 * (sepa - sepa_prev) : 0.0;
        sepa_prev = sepa;
    }
}

template <int PTS>
__device__ double lifetime_wave_pre(const LifeArgs& a, const LifePre<PTS>& P, double norm_log10) {
    const double norm = pow(10.0, norm_log10);
    const int nst = (a.s_begin < a.s_end) ? a.s_end - a.s_begin : 0;
    double part = 0.0;
    double dadt_left = -norm * P.p1[0] / P.p2[0] + P.gw[0];
#pragma unroll
    for (int i = 1; i < PTS; i++) {
        const double dadt_right = -norm * P.p1[i] / P.p2[i] + P.gw[i];
        if (i <= nst) part += P.num[i] / (dadt_left + dadt_right);
        dadt_left = dadt_right;
    }
    return wave_sum(part) - a.target;
}

// PTS: step boundaries per lane whose factors are kept in registers (PTS - 1 steps), 0 for none
template <int PTS>
__global__ void __launch_bounds__(256, (PTS == 6) ? 4 : 1) k_find_2pwl_norm(EvoConsts K, double target_time, const double* __restrict__ mtot,
                                                        const double* __restrict__ mrat, int64_t n, double sepa_init,
                                                        double rchar, double g_in, double g_out, int nsteps,
                                                        double* __restrict__ norm_log10, double* __restrict__ norm_out) {
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (item >= n) return;   // uniform per wave

    LifeArgs a;
    a.mt = mtot[item];
    a.mr = mrat[item];
    a.target = target_time;
    a.rchar = rchar;
    a.g_in = g_in;
    a.g_out = g_out;
    double risco_log10 = log10(3.0 * MY_SCHW * a.mt);
    double x0 = log10(sepa_init);
    a.dx = (x0 - risco_log10) / nsteps;
    int chunk = (nsteps + 63) / 64;
    a.s_begin = min(lane * chunk, nsteps);
    a.s_end = min(a.s_begin + chunk, nsteps);
    double x = x0;
    for (int s = 0; s < a.s_begin; s++) x -= a.dx;
    a.x_lane0 = x;

    // the normalisation-free factors of every step boundary once (registers), when they fit
    LifePre<(PTS > 0) ? PTS : 1> P;
    if constexpr (PTS > 0) lifetime_pre(K, a, P);
    auto life = [&](double v) {
        if constexpr (PTS > 0) return lifetime_wave_pre(a, P, v);
        else return lifetime_wave(K, a, v);
    };

    const double xtol = 1e-3, rtol = 1e-5;
    const int maxiter = 100;
    double xpre = -20.0, xcur = 20.0, xblk = 0.0, fblk = 0.0, spre = 0.0, scur = 0.0;
    double fpre = life(xpre);
    double fcur = life(xcur);
    double res = xcur;
    bool done = false;
    if (fpre == 0) { res = xpre; done = true; }
    else if (fcur == 0) { res = xcur; done = true; }
    else if (signbit(fpre) == signbit(fcur)) { res = 0.0; done = true; }
    for (int it = 0; it < maxiter && !done; it++) {
        if (fpre != 0 && fcur != 0 && (signbit(fpre) != signbit(fcur))) {
            xblk = xpre;
            fblk = fpre;
            spre = scur = xcur - xpre;
        }
        if (fabs(fblk) < fabs(fcur)) {
            xpre = xcur; xcur = xblk; xblk = xpre;
            fpre = fcur; fcur = fblk; fblk = fpre;
        }
        double delta = (xtol + rtol * fabs(xcur)) / 2;
        double sbis = (xblk - xcur) / 2;
        if (fcur == 0 || fabs(sbis) < delta) { res = xcur; done = true; break; }
        if (fabs(spre) > delta && fabs(fcur) < fabs(fpre)) {
            double stry;
            if (xpre == xblk) {
                stry = -fcur * (xcur - xpre) / (fcur - fpre);
            } else {
                double dpre = (fpre - fcur) / (xpre - xcur);
                double dblk = (fblk - fcur) / (xblk - xcur);
                stry = -fcur * (fblk * dblk - fpre * dpre) / (dblk * dpre * (fblk - fpre));
            }
            if (2 * fabs(stry) < fmin(fabs(spre), 3 * fabs(sbis) - delta)) {
                spre = scur;
                scur = stry;
            } else {
                spre = sbis;
                scur = sbis;
            }
        } else {
            spre = sbis;
            scur = sbis;
        }
        xpre = xcur;
        fpre = fcur;
        if (fabs(scur) > delta) xcur += scur;
        else xcur += (sbis > 0 ? delta : -delta);
        fcur = life(xcur);
        res = xcur;
    }
    if (lane == 0) {
        norm_log10[item] = res;
        if (norm_out) norm_out[item] = pow(10.0, res);
    }
}

// Lifetime from sepa_init to the ISCO at a given normalisation (no root finding), one wave per
// element: integrate_binary_evolution_2pwl (sam_cyutils.pyx:401-418).
__global__ void __launch_bounds__(256) k_lifetime_2pwl(EvoConsts K, const double* __restrict__ norm_log10,
                                                       const double* __restrict__ mtot, const double* __restrict__ mrat,
                                                       int64_t n, double sepa_init, double rchar, double g_in,
                                                       double g_out, int nsteps, double* __restrict__ time) {
    const int lane = threadIdx.x & 63;
    const int64_t item = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (item >= n) return;   // uniform per wave
    LifeArgs a;
    a.mt = mtot[item];
    a.mr = mrat[item];
    a.target = 0.0;
    a.rchar = rchar;
    a.g_in = g_in;
    a.g_out = g_out;
    const double x0 = log10(sepa_init);
    a.dx = (x0 - log10(3.0 * MY_SCHW * a.mt)) / nsteps;
    const int chunk = (nsteps + 63) / 64;
    a.s_begin = min(lane * chunk, nsteps);
    a.s_end = min(a.s_begin + chunk, nsteps);
    double x = x0;
    for (int s = 0; s < a.s_begin; s++) x -= a.dx;
    a.x_lane0 = x;
    const double t = lifetime_wave(K, a, norm_log10[item]);
    if (lane == 0) time[item] = t;
}

__global__ void k_hard_func_2pwl_gw(EvoConsts K, const double* mtot, const double* mrat, const double* sepa,
                                    const double* norm, int64_t n, double rchar, double g_in, double g_out,
                                    double* dadt) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dadt[i] = hard_func_2pwl_gw(K, mtot[i], mrat[i], sepa[i], norm[i], rchar, g_in, g_out);
}

// ---------------------------------------------------------------------------
// GW-only dynamic binary number: one thread per (M,Q,Z) corner, looping F.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_dbn_gw(EvoConsts K, const double* __restrict__ fobs, int F,
                                                const double* __restrict__ nden, const double* __restrict__ mtot, int M,
                                                const double* __restrict__ mrat, int Q, int Z,
                                                const double* __restrict__ redz_prime,
                                                const double* __restrict__ gz, const double* __restrict__ gdc, int n_interp,
                                                double* __restrict__ redz_final, double* __restrict__ diff_num) {
    int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= (int64_t)M * Q * Z) return;
    int ii = (int)(cell / ((int64_t)Q * Z));
    int jj = (int)((cell / Z) % Q);
    double* rf = redz_final + cell * F;
    double* dn = diff_num + cell * F;
    for (int ff = 0; ff < F; ff++) {
        rf[ff] = -1.0;
        dn[ff] = 0.0;
    }
    double rzp = redz_prime[cell];
    if (rzp <= 0.0) return;
    double mt = mtot[ii], mr = mrat[jj];
    double rad_isco = 3.0 * MY_SCHW * mt;
    double frst_isco = kepler_freq_from_sepa(K, mt, rad_isco);
    int idx = ww_decreasing(0, n_interp, rzp, gz);
    double dcom = interp_at_index(idx, rzp, gz, gdc);
    double dmpc = dcom / MY_MPC;
    double cosmo_fact = K.four_pi_c_mpc * (1.0 + rzp) * (dmpc * dmpc);
    double nd = nden[cell];
    for (int ff = 0; ff < F; ff++) {
        rf[ff] = rzp;
        double ft = fobs[ff];
        double frst = ft * (1.0 + rzp);
        if (frst > frst_isco) {
            for (int fp = ff + 1; fp < F; fp++) rf[fp] = rzp;
            break;
        }
        double sepa = kepler_sepa_from_freq(K, mt, frst);
        double dadt = hard_gw(K, mt, mr, sepa);
        double tres = -(2.0 / 3.0) * sepa / dadt;
        dn[ff] = nd * tres * cosmo_fact;
    }
}

// ---------------------------------------------------------------------------
// 2PL dynamic binary number.
// One workgroup per (M,Q) pair.  The workgroup first tabulates the evolution
// steps in LDS (separation chain, rest-frame frequency, dt, cumulative time;
// the chain and the cumulative sum are sequential exactly as in the
// reference).  Then
//   phase 1: a few threads per redshift corner kk each walk a contiguous share
//            of the steps, bracketing the target frequencies; every target
//            keeps the LARGEST step that brackets it (an atomic max: the
//            reference's last write over its in-order steps, per (kk, ff)),
//            honouring the reference's break when the interpolated time passes
//            the age of the universe.  The steps are kept in LDS (or, when
//            Z x F does not fit, in redz_final, which the workgroup owns);
//   phase 2: one thread per (kk, ff) recomputes that step's bracket and does
//            the per-target work (Kepler separation, hardening rate, residence
//            time, cosmological factor) with every lane busy.
// Table walks: the reference walks from the previous index (while_while_
// increasing); the times walked are non-decreasing along the steps, so a walk
// from the start of the table gives the same interval (the binary searches in
// phase 2 compute exactly that interval).
// ---------------------------------------------------------------------------
constexpr int DBN_MAX_STEPS = 4096;
constexpr int DBN_THREADS = 256;

// first index i in [0, n-2] with edges[i+1] >= val (n-2 if none): the
// interval a forward walk of while_while_increasing stops at
__device__ __forceinline__ int bracket_increasing(int n, double val, const double* edges) {
    int lo = 0, hi = n - 2;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (edges[mid + 1] < val) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

constexpr int DBN_FOBS_LDS = 1024;   // target frequencies staged in LDS up to this many

template <bool RF_LDS>
__global__ void __launch_bounds__(DBN_THREADS, 5) k_dbn_2pwl(EvoConsts K, const double* __restrict__ fobs, int F,
                                                          int sorted_f, double sepa_init, int nsteps,
                                                          const double* __restrict__ hard_norm, double rchar,
                                                          double g_in, double g_out, const double* __restrict__ nden,
                                                          const double* __restrict__ mtot, int M,
                                                          const double* __restrict__ mrat, int Q,
                                                          const double* __restrict__ redz, int Z,
                                                          const double* __restrict__ gmt_time,
                                                          const double* __restrict__ gz_g, const double* __restrict__ gdc_g,
                                                          const double* __restrict__ gage_g, int n_interp,
                                                          double* __restrict__ redz_final, double* __restrict__ diff_num) {
    extern __shared__ double dbn_lds[];
    const int ns1 = nsteps + 1;
    double* s_frst = dbn_lds;
    double* s_dt = s_frst + ns1;
    double* s_tevo = s_dt + ns1;
    double* s_gz = s_tevo + ns1;
    double* s_gdc = s_gz + n_interp;
    double* s_gage = s_gdc + n_interp;
    double* s_agek = s_gage + n_interp;   // (Z)
    double* s_fobs = s_agek + Z;          // (F) when F <= DBN_FOBS_LDS
    // the set-up's separations and rates, then (their last reader is past a barrier before the
    // recorded steps are cleared) the (Z, F) recorded steps when RF_LDS, in the same space
    double* s_sepa = s_fobs + (F <= DBN_FOBS_LDS ? F : 0);
    double* s_dadt = s_sepa + ns1;
    int* s_rf = (int*)s_sepa;

    const int ii = blockIdx.x / Q;
    const int jj = blockIdx.x % Q;
    const int tid = threadIdx.x;
    const double mt = mtot[ii], mr = mrat[jj];
    const double norm = hard_norm[(int64_t)ii * Q + jj];
    const int64_t row0 = ((int64_t)ii * Q + jj) * Z;     // first corner of this (M,Q) column
    double* rf0 = redz_final + row0 * F;
    double* dn0 = diff_num + row0 * F;

    for (int i = tid; i < n_interp; i += blockDim.x) {
        s_gz[i] = gz_g[i];
        s_gdc[i] = gdc_g[i];
        s_gage[i] = gage_g[i];
    }
    // the target frequencies are searched at every evolution step: from LDS, not global memory
    const bool fobs_lds = F <= DBN_FOBS_LDS;
    if (fobs_lds)
        for (int i = tid; i < F; i += blockDim.x) s_fobs[i] = fobs[i];
    const double* fo = fobs_lds ? s_fobs : fobs;
    // separation chain (sequential, as sam_cyutils.pyx:611-637)
    if (tid == 0) {
        double x = log10(sepa_init);
        double risco = 3.0 * MY_SCHW * mt;
        double dx = (x - log10(risco)) / nsteps;
        s_sepa[0] = pow(10.0, x);
        for (int s = 1; s <= nsteps; s++) {
            x -= dx;
            s_sepa[s] = x;   // log10 for now
        }
    }
    __syncthreads();
    for (int s = tid; s <= nsteps; s += blockDim.x) {
        double sp = (s == 0) ? s_sepa[0] : pow(10.0, s_sepa[s]);
        s_frst[s] = kepler_freq_from_sepa(K, mt, sp);
        s_dadt[s] = hard_func_2pwl_gw(K, mt, mr, sp, norm, rchar, g_in, g_out);
        s_sepa[s] = sp;   // each s is read and written by the same thread
    }
    // age at each redshift edge from the table (sam_cyutils.pyx:586-601): the reference walks the
    // decreasing redshift nodes from the start while gz[ia + 1] > zk; the same index by a binary
    // search for the first node at or below zk
    for (int kk = tid; kk < Z; kk += blockDim.x) {
        const double zk = redz[kk];
        int lo = 1, hi = n_interp;   // first j in [1, n) with gz[j] <= zk, n if none
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_gz[mid] > zk) lo = mid + 1;
            else hi = mid;
        }
        int ia = lo - 1;
        if (ia > n_interp - 2) ia = n_interp - 2;
        s_agek[kk] = interp_between(zk, s_gz[ia], s_gz[ia + 1], s_gage[ia], s_gage[ia + 1]);
    }
    __syncthreads();
    for (int s = tid + 1; s <= nsteps; s += blockDim.x)
        s_dt[s] = 2.0 * (s_sepa[s] - s_sepa[s - 1]) / (s_dadt[s - 1] + s_dadt[s]);
    __syncthreads();
    if (tid == 0) {   // the reference's sequential sum; unrolled so that the loads run ahead of it
        double t = 0.0;
        s_tevo[0] = 0.0;
#pragma unroll 8
        for (int s = 1; s <= nsteps; s++) {
            t += s_dt[s];
            s_tevo[s] = t;
        }
    }
    __syncthreads();

    const double age_universe = s_gage[n_interp - 1];
    // the step recorded per (kk, ff): the largest that brackets the target, as the reference's
    // last write over its in-order steps; -1 none.  In LDS when it fits, else in redz_final as
    // 64-bit step + 1 (0 none)
    const int nzf = Z * F;
    unsigned long long* rfg = (unsigned long long*)rf0;
    for (int p = tid; p < nzf; p += blockDim.x) {
        if (RF_LDS) s_rf[p] = -1;
        else rfg[p] = 0ull;
    }
    __syncthreads();
    // phase 1: S threads per redshift corner kk, each walking a contiguous share of the steps
    // (its first table interval by a binary search: the interval a walk from the start reaches)
    const int S = max(1, (int)blockDim.x / Z);
    for (int t = tid; t < Z * S; t += blockDim.x) {
        const int kk = t / S, seg = t - kk * S;
        const int st0 = (int)((int64_t)nsteps * seg / S), st1 = (int)((int64_t)nsteps * (seg + 1) / S);
        const double age_k = s_agek[kk];
        const double gmt = gmt_time[row0 + kk];
        int idx_left = 0, ff_lo = 0;
        if (seg > 0) idx_left = bracket_increasing(n_interp, s_tevo[st0 + 1] + gmt + age_k - s_dt[st0 + 1], s_gage);
        // The nodes of the last walks' intervals are kept in registers: a walk that would not move
        // (the time inside the cached interval) reads nothing.  Left: the reference's walk from the
        // previous interval stays put exactly when gage[il] <= t <= gage[il + 1].  Right: the walk
        // from idx_left ends in the left interval when t <= gage[idx_left + 1], in the cached right
        // interval ir >= idx_left when gage[ir] < t <= gage[ir + 1].  Targets: the first target
        // >= fl stays put when fo[ff_lo - 1] < fl <= fo[ff_lo].
        int il = -1, ir = -1;
        double gl0 = 0.0, gl1 = 0.0, zl0 = 0.0, zl1 = 0.0, gr0 = 0.0, gr1 = 0.0, zr0 = 0.0, zr1 = 0.0;
        double f_below = -1.0, f_at = -1.0;   // fo[ff_lo - 1], fo[ff_lo] (+-inf past the ends); unset
        bool f_set = false;
        for (int step = st0; step < st1; step++) {
            const double dt = s_dt[step + 1];
            const double time_right = s_tevo[step + 1] + gmt + age_k;
            const double time_left = time_right - dt;
            if (time_left > age_universe) continue;
            if (!(il >= 0 && gl0 <= time_left && time_left <= gl1)) {
                idx_left = ww_increasing(idx_left, n_interp, time_left, s_gage);
                il = idx_left;
                gl0 = s_gage[il];
                gl1 = s_gage[il + 1];
                zl0 = s_gz[il];
                zl1 = s_gz[il + 1];
            }
            const double redz_left = interp_between(time_left, gl0, gl1, zl0, zl1);
            if (redz_left < 0.0) continue;
            double redz_right;
            if (time_right <= gl1) {
                redz_right = interp_between(time_right, gl0, gl1, zl0, zl1);
            } else {
                if (!(ir >= idx_left && gr0 < time_right && time_right <= gr1)) {
                    ir = ww_increasing(idx_left, n_interp, time_right, s_gage);
                    gr0 = s_gage[ir];
                    gr1 = s_gage[ir + 1];
                    zr0 = s_gz[ir];
                    zr1 = s_gz[ir + 1];
                }
                redz_right = interp_between(time_right, gr0, gr1, zr0, zr1);
            }
            if (redz_right < 0.0) redz_right = 0.0;
            const double fl = s_frst[step] / (1.0 + redz_left);
            const double fr = s_frst[step + 1] / (1.0 + redz_right);
            if (sorted_f) {
                // first target >= fl (targets are non-decreasing), moved from the previous step's:
                // fl grows along the evolution, so it moves by a few places at most, and the two
                // walks give the first such target for any fl
                if (!(f_set && f_below < fl && fl <= f_at)) {
                    while (ff_lo > 0 && fo[ff_lo - 1] >= fl) ff_lo--;
                    while (ff_lo < F && fo[ff_lo] < fl) ff_lo++;
                    f_below = (ff_lo > 0) ? fo[ff_lo - 1] : -HUGE_VAL;
                    f_at = (ff_lo < F) ? fo[ff_lo] : HUGE_VAL;
                    f_set = true;
                }
                if (!(f_at <= fr)) continue;   // no target in [fl, fr]
                for (int ff = ff_lo; ff < F; ff++) {
                    const double ft = fo[ff];
                    if (fr < ft) break;   // past the contiguous bracketed range
                    const double new_time = interp_between(ft, fl, fr, time_left, time_right);
                    if (new_time > age_universe) break;
                    if (RF_LDS) atomicMax(s_rf + kk * F + ff, step);
                    else atomicMax(rfg + (int64_t)kk * F + ff, (unsigned long long)step + 1ull);
                }
            } else {
                for (int ff = 0; ff < F; ff++) {
                    const double ft = fo[ff];
                    if ((ft < fl) || (fr < ft)) continue;
                    const double new_time = interp_between(ft, fl, fr, time_left, time_right);
                    if (new_time > age_universe) break;
                    if (RF_LDS) atomicMax(s_rf + kk * F + ff, step);
                    else atomicMax(rfg + (int64_t)kk * F + ff, (unsigned long long)step + 1ull);
                }
            }
        }
    }
    __syncthreads();   // every recorded step is in place
    const double mt13 = pow(mt, 1.0 / 3.0);
    // phase 2: the per-target work for every recorded (kk, ff)
    for (int p = tid; p < nzf; p += blockDim.x) {
        const int step = RF_LDS ? s_rf[p] : (int)((long long)rfg[p] - 1);
        if (step < 0) {
            rf0[p] = -1.0;
            dn0[p] = 0.0;
            continue;
        }
        const int kk = p / F, ff = p - (p / F) * F;
        const double gmt = gmt_time[row0 + kk];
        const double nd = nden[row0 + kk];
        const double dt = s_dt[step + 1];
        const double time_right = s_tevo[step + 1] + gmt + s_agek[kk];
        const double time_left = time_right - dt;
        const int idx_left = bracket_increasing(n_interp, time_left, s_gage);
        const double redz_left = interp_at_index(idx_left, time_left, s_gage, s_gz);
        const int idx_right = bracket_increasing(n_interp, time_right, s_gage);
        double redz_right = interp_at_index(idx_right, time_right, s_gage, s_gz);
        if (redz_right < 0.0) redz_right = 0.0;
        const double fl = s_frst[step] / (1.0 + redz_left);
        const double fr = s_frst[step + 1] / (1.0 + redz_right);
        const double ft = fo[ff];
        const double new_time = interp_between(ft, fl, fr, time_left, time_right);
        const int inew = ww_increasing(idx_left, n_interp, new_time, s_gage);
        const double new_redz = interp_at_index(inew, new_time, s_gage, s_gz);
        const double dcom = interp_at_index(inew, new_time, s_gage, s_gdc);
        rf0[p] = new_redz;
        const double frst = ft * (1.0 + new_redz);
        // Kepler separation and the 2PL + GW rate at the target (sam_cyutils.pyx:240-253, 52-61)
        // with the bin-independent factor hoisted and cheaper powers: f^(2/3) as cbrt(f^2), the
        // power laws of half-integer exponents from products and one square root (a few ulp)
        const double sepa = K.kepler_sepa * mt13 / cbrt(frst * frst);
        const double xx = sepa / rchar;
        const double dadt = -norm * pow_halfint(1.0 + xx, g_in - g_out) / pow_halfint(xx, g_in - 1.0) +
                            hard_gw(K, mt, mr, sepa);
        const double tres = -(2.0 / 3.0) * sepa / dadt;
        const double dmpc = dcom / MY_MPC;
        const double cosmo_fact = K.four_pi_c_mpc * (1.0 + new_redz) * (dmpc * dmpc);
        dn0[p] = nd * tres * cosmo_fact;
    }
}

// ---------------------------------------------------------------------------
// 3D trapezoid x 1D Riemann integration (sam_cyutils.pyx:166-216).
// One thread per output (m,q,z,f); corner order and operation order follow
// the reference loop.
// ---------------------------------------------------------------------------
// block b: (mass, ratio) bin b / nchunk, chunk b % nchunk of its 256-output chunks of (redshift
// bin, frequency) -- the chunks of a bin consecutive, as the corner rows they share; a thread
// decodes its output with one 32-bit division (the flat decode took three 64-bit ones)
__global__ void __launch_bounds__(256) k_integrate_3dx1d(const double* __restrict__ lm, int M, const double* __restrict__ mr,
                                                         int Q, const double* __restrict__ rz, int Z,
                                                         const double* __restrict__ dlnf, int F,
                                                         const double* __restrict__ dnum, double* __restrict__ numb) {
    const uint32_t per_mq = (uint32_t)(Z - 1) * (uint32_t)F;
    const uint32_t nchunk = (per_mq + blockDim.x - 1) / blockDim.x;
    const int mq = (int)(blockIdx.x / nchunk);   // uniform: the block's chunks of one bin are consecutive
    const int mm = mq / (Q - 1), qq = mq - mm * (Q - 1);
    const uint32_t loc = (blockIdx.x - (uint32_t)mq * nchunk) * blockDim.x + threadIdx.x;
    if (loc >= per_mq) return;
    const int zz = (int)(loc / (uint32_t)F), ff = (int)(loc - (uint32_t)zz * (uint32_t)F);
    const int64_t o = (int64_t)mq * per_mq + loc;
    double dm = lm[mm + 1] - lm[mm];
    double dmdq = dm * (mr[qq + 1] - mr[qq]);
    double dmdqdz = dmdq * (rz[zz + 1] - rz[zz]);
    double temp = 0.0;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int k = 0; k < 2; k++)
                temp += dnum[((((int64_t)(mm + i) * Q) + (qq + j)) * Z + (zz + k)) * F + ff];
    numb[o] = temp * dmdqdz * dlnf[ff] / 8.0;
}

}  // namespace hd

using namespace hd;

extern "C" int hd_find_2pwl_norm(double target_time, const double* mtot, const double* mrat, int64_t n,
                                 double sepa_init, double rchar, double gamma_inner, double gamma_outer, int nsteps,
                                 double* norm_log10, double* norm, void* stream) {
    if (n < 0 || nsteps < 1 || !mtot || !mrat || !norm_log10) return set_error(HD_ERR_ARG, "hd_find_2pwl_norm: bad arguments");
    if (n == 0) return HD_OK;
    const int waves_per_block = 4;
    dim3 grid((unsigned)cdiv(n, waves_per_block));
    const int chunk = (nsteps + 63) / 64;   // steps per lane
    if (chunk <= 5)
        hipLaunchKernelGGL(k_find_2pwl_norm<6>, grid, dim3(64 * waves_per_block), 0, (hipStream_t)stream, evo_consts(),
                           target_time, mtot, mrat, n, sepa_init, rchar, gamma_inner, gamma_outer, nsteps, norm_log10, norm);
    else if (chunk <= 8)
        hipLaunchKernelGGL(k_find_2pwl_norm<9>, grid, dim3(64 * waves_per_block), 0, (hipStream_t)stream, evo_consts(),
                           target_time, mtot, mrat, n, sepa_init, rchar, gamma_inner, gamma_outer, nsteps, norm_log10, norm);
    else
        hipLaunchKernelGGL(k_find_2pwl_norm<0>, grid, dim3(64 * waves_per_block), 0, (hipStream_t)stream, evo_consts(),
                           target_time, mtot, mrat, n, sepa_init, rchar, gamma_inner, gamma_outer, nsteps, norm_log10, norm);
    return check_launch("k_find_2pwl_norm");
}

extern "C" int hd_hard_func_2pwl_gw(const double* mtot, const double* mrat, const double* sepa, const double* norm,
                                    int64_t n, double rchar, double gamma_inner, double gamma_outer, double* dadt,
                                    void* stream) {
    if (n < 0 || (n > 0 && (!mtot || !mrat || !sepa || !norm || !dadt)))
        return set_error(HD_ERR_ARG, "hd_hard_func_2pwl_gw: bad arguments");
    if (n == 0) return HD_OK;
    hipLaunchKernelGGL(k_hard_func_2pwl_gw, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       evo_consts(), mtot, mrat, sepa, norm, n, rchar, gamma_inner, gamma_outer, dadt);
    return check_launch("k_hard_func_2pwl_gw");
}

extern "C" int hd_dbn_gw(const double* fobs_orb, int F, const double* nden, const double* mtot, int M,
                         const double* mrat, int Q, int Z, const double* redz_prime, const double* grid_z,
                         const double* grid_dcom, int n_interp, double* redz_final, double* diff_num, void* stream) {
    if (F < 1 || M < 1 || Q < 1 || Z < 1 || n_interp < 2)
        return set_error(HD_ERR_ARG, "hd_dbn_gw: bad sizes F=%d M=%d Q=%d Z=%d n_interp=%d", F, M, Q, Z, n_interp);
    int64_t n = (int64_t)M * Q * Z;
    hipLaunchKernelGGL(k_dbn_gw, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, evo_consts(), fobs_orb,
                       F, nden, mtot, M, mrat, Q, Z, redz_prime, grid_z, grid_dcom, n_interp, redz_final, diff_num);
    return check_launch("k_dbn_gw");
}

extern "C" int hd_dbn_2pwl(const double* fobs_orb, int F, int fobs_sorted, double sepa_init, int num_steps,
                           const double* hard_norm, double rchar, double gamma_inner, double gamma_outer,
                           const double* nden, const double* mtot, int M, const double* mrat, int Q, const double* redz,
                           int Z, const double* gmt_time, const double* grid_z, const double* grid_dcom,
                           const double* grid_age, int n_interp, double* redz_final, double* diff_num, void* stream) {
    if (F < 1 || M < 1 || Q < 1 || Z < 1 || n_interp < 2 || num_steps < 1)
        return set_error(HD_ERR_ARG, "hd_dbn_2pwl: bad sizes F=%d M=%d Q=%d Z=%d n_interp=%d steps=%d", F, M, Q, Z,
                         n_interp, num_steps);
    if (num_steps > DBN_MAX_STEPS) return set_error(HD_ERR_SIZE, "hd_dbn_2pwl: num_steps=%d > %d", num_steps, DBN_MAX_STEPS);
    if ((int64_t)M * Q > 0x7fffffff) return set_error(HD_ERR_SIZE, "hd_dbn_2pwl: M*Q too large");
    if ((int64_t)Z * F > 0x7fffffff) return set_error(HD_ERR_SIZE, "hd_dbn_2pwl: Z*F too large");
    const size_t lds = ((size_t)5 * (num_steps + 1) + 3 * (size_t)n_interp + (size_t)Z +
                        (F <= DBN_FOBS_LDS ? (size_t)F : 0)) * sizeof(double);
    if (lds > 65536)
        return set_error(HD_ERR_SIZE, "hd_dbn_2pwl: step/cosmology tables need %zu B of LDS (> 64 KiB)", lds);
    // with the recorded steps, which share the set-up's separations and rates (2 (num_steps + 1) doubles)
    const size_t lds_rf = lds - (size_t)2 * (num_steps + 1) * sizeof(double) +
                          std::max((size_t)2 * (num_steps + 1) * sizeof(double), (size_t)Z * F * sizeof(int));
    if (lds_rf <= 65536)
        hipLaunchKernelGGL(k_dbn_2pwl<true>, dim3(M * Q), dim3(DBN_THREADS), lds_rf, (hipStream_t)stream, evo_consts(),
                           fobs_orb, F, fobs_sorted, sepa_init, num_steps, hard_norm, rchar, gamma_inner, gamma_outer,
                           nden, mtot, M, mrat, Q, redz, Z, gmt_time, grid_z, grid_dcom, grid_age, n_interp,
                           redz_final, diff_num);
    else
        hipLaunchKernelGGL(k_dbn_2pwl<false>, dim3(M * Q), dim3(DBN_THREADS), lds, (hipStream_t)stream, evo_consts(),
                           fobs_orb, F, fobs_sorted, sepa_init, num_steps, hard_norm, rchar, gamma_inner, gamma_outer,
                           nden, mtot, M, mrat, Q, redz, Z, gmt_time, grid_z, grid_dcom, grid_age, n_interp,
                           redz_final, diff_num);
    return check_launch("k_dbn_2pwl");
}

extern "C" int hd_lifetime_2pwl(const double* norm_log10, const double* mtot, const double* mrat, int64_t n,
                                double sepa_init, double rchar, double gamma_inner, double gamma_outer, int nsteps,
                                double* time, void* stream) {
    if (n < 0 || nsteps < 1 || (n > 0 && (!norm_log10 || !mtot || !mrat || !time)))
        return set_error(HD_ERR_ARG, "hd_lifetime_2pwl: bad arguments");
    if (n == 0) return HD_OK;
    hipLaunchKernelGGL(k_lifetime_2pwl, dim3((unsigned)cdiv(n, 4)), dim3(256), 0, (hipStream_t)stream, evo_consts(),
                       norm_log10, mtot, mrat, n, sepa_init, rchar, gamma_inner, gamma_outer, nsteps, time);
    return check_launch("k_lifetime_2pwl");
}

extern "C" int hd_integrate_3dx1d(const double* log10_mtot, int M, const double* mrat, int Q, const double* redz, int Z,
                                  const double* dln_freq, int F, const double* dnum, double* numb, void* stream) {
    if (M < 2 || Q < 2 || Z < 2 || F < 1) return set_error(HD_ERR_ARG, "hd_integrate_3dx1d: bad sizes");
    const int64_t per_mq = (int64_t)(Z - 1) * F;
    const int64_t nblk = (int64_t)(M - 1) * (Q - 1) * cdiv(per_mq, 256);
    if (per_mq > 0xffffffffll || nblk > 0x7fffffffll) return set_error(HD_ERR_SIZE, "hd_integrate_3dx1d: grid too large");
    hipLaunchKernelGGL(k_integrate_3dx1d, dim3((unsigned)nblk), dim3(256), 0,
                       (hipStream_t)stream, log10_mtot, M, mrat, Q, redz, Z, dln_freq, F, dnum, numb);
    return check_launch("k_integrate_3dx1d");
}
