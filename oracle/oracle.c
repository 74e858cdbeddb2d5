/*
 * oracle.c -- FP64 CPU restatement of holodeck's semi-analytic GWB hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (holodeck_amd/) may link,
 * load or call this file.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg as
 * the timed CPU baseline.
 *
 * Every function follows the reference loop nest literally (same iteration
 * order, same operation order, same early exits, same -1 sentinels), so that
 * the GPU kernels can be compared against it at FP64 rounding level.
 * Citations are to the reference sources (holodeck v1.6).
 *
 * Pinning (see tests/): GW/Kepler constants and formulas against the golden
 * vectors of holodeck/tests/test_utils.py:60-94; brentq against
 * scipy.optimize.brentq (the same published Brent algorithm the reference
 * calls through scipy.optimize.cython_optimize); the dynamic-binary-number
 * loops against brute-force restatements of the reference's own Python
 * versions (sam.py:468-676) and closed forms; the realization walks against
 * brute-force numpy with injected counts.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* ---- constants: holodeck/sams/sam_cyutils.pyx:30-41, holodeck/cyutils.pyx:39-48 ---- */
static const double MY_NWTG = 6.6742999e-08;
static const double MY_SPLC = 29979245800.0;
static const double MY_MPC = 3.08567758e+24;
static const double MY_SCHW = 1.4852320538237328e-28;

static double GW_DADT_SEP_CONST;
static double KEPLER_CONST_FREQ;
static double KEPLER_CONST_SEPA;
static double FOUR_PI_SPLC_OVER_MPC;
static int g_init = 0;

static void or_init(void) {
    if (g_init) return;
    /* sam_cyutils.pyx:36-41, evaluated left to right as C does */
    GW_DADT_SEP_CONST = -64.0 * pow(MY_NWTG, 3) / 5.0 / pow(MY_SPLC, 5);
    KEPLER_CONST_FREQ = (1.0 / (2.0 * M_PI)) * sqrt(MY_NWTG);
    KEPLER_CONST_SEPA = pow(MY_NWTG, 1.0 / 3.0) / pow(2.0 * M_PI, 2.0 / 3.0);
    FOUR_PI_SPLC_OVER_MPC = 4 * M_PI * MY_SPLC / MY_MPC;
    g_init = 1;
}

void or_constants(double *out) {
    or_init();
    out[0] = GW_DADT_SEP_CONST;
    out[1] = KEPLER_CONST_FREQ;
    out[2] = KEPLER_CONST_SEPA;
    out[3] = FOUR_PI_SPLC_OVER_MPC;
    out[4] = MY_SCHW;
    out[5] = MY_MPC;
}

/* sam_cyutils.pyx:45-49 */
double or_hard_gw(double mtot, double mrat, double sepa) {
    or_init();
    return GW_DADT_SEP_CONST * pow(mtot, 3) * mrat / pow(sepa, 3) / pow(1 + mrat, 2);
}

/* sam_cyutils.pyx:52-55 */
static double kepler_freq_from_sepa(double mtot, double sepa) {
    return KEPLER_CONST_FREQ * sqrt(mtot) / pow(sepa, 1.5);
}

/* sam_cyutils.pyx:58-61 */
static double kepler_sepa_from_freq(double mtot, double freq) {
    return KEPLER_CONST_SEPA * pow(mtot, 1.0 / 3.0) / pow(freq, 2.0 / 3.0);
}

/* sam_cyutils.pyx:64-84 (the `ii+1 < size` guard only avoids the reference's
 * unreachable out-of-bounds read; results are unchanged for valid inputs) */
static int ww_increasing(int start, int size, double val, const double *edges) {
    int index = start;
    while ((index < size - 2) && (edges[index + 1] < val)) index += 1;
    while ((index > 0) && (edges[index] > val)) index -= 1;
    return index;
}

/* sam_cyutils.pyx:87-107 */
static int ww_decreasing(int start, int size, double val, const double *edges) {
    int index = start;
    while ((index < size - 1) && (edges[index + 1] > val)) index += 1;
    while ((index > 0) && (edges[index - 1] < val)) index -= 1;
    return index;
}

/* cyutils.pyx:322-327 */
static double interp_between(double xnew, double xl, double xr, double yl, double yr) {
    return yl + (yr - yl) * (xnew - xl) / (xr - xl);
}

/* cyutils.pyx:330-358 */
static double interp_at_index(int idx, double xnew, const double *xold, const double *yold) {
    return interp_between(xnew, xold[idx], xold[idx + 1], yold[idx], yold[idx + 1]);
}

/* ---------------------------------------------------------------------------
 * integrate_differential_number_3dx1d  (sam_cyutils.pyx:115-216)
 * log10_mtot (M), mrat (Q), redz (Z), dln_freq (F); dnum (M,Q,Z,F) -> numb (M-1,Q-1,Z-1,F)
 * ------------------------------------------------------------------------- */
void or_integrate_3dx1d(const double *log10_mtot, int n_mtot, const double *mrat, int n_mrat,
                        const double *redz, int n_redz, const double *dln_freq, int n_freq,
                        const double *dnum, double *numb) {
    for (int mm = 0; mm < n_mtot - 1; mm++) {
        double dm = log10_mtot[mm + 1] - log10_mtot[mm];
        for (int qq = 0; qq < n_mrat - 1; qq++) {
            double dmdq = dm * (mrat[qq + 1] - mrat[qq]);
            for (int zz = 0; zz < n_redz - 1; zz++) {
                double dmdqdz = dmdq * (redz[zz + 1] - redz[zz]);
                for (int ff = 0; ff < n_freq; ff++) {
                    double temp = 0.0;
                    for (int ii = 0; ii < 2; ii++)
                        for (int jj = 0; jj < 2; jj++)
                            for (int kk = 0; kk < 2; kk++) {
                                size_t src = (((size_t)(mm + ii) * n_mrat + (qq + jj)) * n_redz + (zz + kk)) * n_freq + ff;
                                temp += dnum[src];
                            }
                    size_t dst = (((size_t)mm * (n_mrat - 1) + qq) * (n_redz - 1) + zz) * n_freq + ff;
                    numb[dst] = temp * dmdqdz * dln_freq[ff] / 8.0;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Fixed_Time_2PL_SAM hardening  (sam_cyutils.pyx:240-398)
 * ------------------------------------------------------------------------- */
static double hard_func_2pwl(double norm, double xx, double gamma_inner, double gamma_outer) {
    return -norm * pow(1.0 + xx, -gamma_outer + gamma_inner) / pow(xx, gamma_inner - 1);
}

double or_hard_func_2pwl_gw(double mtot, double mrat, double sepa, double norm, double rchar,
                            double gamma_inner, double gamma_outer) {
    or_init();
    double dadt = hard_func_2pwl(norm, sepa / rchar, gamma_inner, gamma_outer);
    dadt += or_hard_gw(mtot, mrat, sepa);
    return dadt;
}

typedef struct {
    double target_time, mt, mr, sepa_init, rchar, gamma_inner, gamma_outer;
    int nsteps;
} lifetime_params;

/* sam_cyutils.pyx:356-398 */
static double lifetime_2pwl(double norm_log10, const lifetime_params *p) {
    double risco_log10 = log10(3.0 * MY_SCHW * p->mt);
    double sepa_log10 = log10(p->sepa_init);
    double norm = pow(10.0, norm_log10);
    double dx = (sepa_log10 - risco_log10) / p->nsteps;
    double time = 0.0;
    double sepa_left = pow(10.0, sepa_log10);
    double dadt_left = or_hard_func_2pwl_gw(p->mt, p->mr, sepa_left, norm, p->rchar, p->gamma_inner, p->gamma_outer);
    for (int ii = 0; ii < p->nsteps; ii++) {
        sepa_log10 -= dx;
        double sepa_right = pow(10.0, sepa_log10);
        double dadt_right = or_hard_func_2pwl_gw(p->mt, p->mr, sepa_right, norm, p->rchar, p->gamma_inner, p->gamma_outer);
        double dt = 2.0 * (sepa_right - sepa_left) / (dadt_left + dadt_right);
        time += dt;
        sepa_left = sepa_right;
        dadt_left = dadt_right;
    }
    return time - p->target_time;
}

double or_lifetime_2pwl(double norm_log10, double mt, double mr, double target_time, double sepa_init,
                        double rchar, double gamma_inner, double gamma_outer, int nsteps) {
    or_init();
    lifetime_params p = {target_time, mt, mr, sepa_init, rchar, gamma_inner, gamma_outer, nsteps};
    return lifetime_2pwl(norm_log10, &p);
}

/*
 * Brent's method (R. P. Brent, "Algorithms for Minimization without
 * Derivatives", 1973, ch. 4), in the bracketing form that scipy's brentq
 * implements and that sam_cyutils.pyx:346-349 calls through
 * scipy.optimize.cython_optimize.brentq.  Returns the last iterate if maxiter
 * is exhausted and 0.0 on a sign error, as that solver does.
 * `iters` (may be NULL) receives the iteration count; `fcalls` the number of
 * function evaluations.
 */
double or_brentq_lifetime(double xa, double xb, double xtol, double rtol, int maxiter,
                          double mt, double mr, double target_time, double sepa_init, double rchar,
                          double gamma_inner, double gamma_outer, int nsteps, int *iters, int *fcalls) {
    or_init();
    lifetime_params p = {target_time, mt, mr, sepa_init, rchar, gamma_inner, gamma_outer, nsteps};
    double xpre = xa, xcur = xb, xblk = 0.0, fblk = 0.0, spre = 0.0, scur = 0.0;
    double fpre = lifetime_2pwl(xpre, &p);
    double fcur = lifetime_2pwl(xcur, &p);
    int nf = 2, it = 0;
    double res;
    if (fpre == 0) { res = xpre; goto done; }
    if (fcur == 0) { res = xcur; goto done; }
    if (signbit(fpre) == signbit(fcur)) { res = 0.0; goto done; }
    for (it = 0; it < maxiter;) {
        it++;
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
        if (fcur == 0 || fabs(sbis) < delta) { res = xcur; goto done; }
        if (fabs(spre) > delta && fabs(fcur) < fabs(fpre)) {
            double stry;
            if (xpre == xblk) {
                stry = -fcur * (xcur - xpre) / (fcur - fpre);
            } else {
                double dpre = (fpre - fcur) / (xpre - xcur);
                double dblk = (fblk - fcur) / (xblk - xcur);
                stry = -fcur * (fblk * dblk - fpre * dpre) / (dblk * dpre * (fblk - fpre));
            }
            double lim = fabs(spre);
            double alt = 3 * fabs(sbis) - delta;
            if (alt < lim) lim = alt;
            if (2 * fabs(stry) < lim) {
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
        fcur = lifetime_2pwl(xcur, &p);
        nf++;
    }
    res = xcur;
done:
    if (iters) *iters = it;
    if (fcalls) *fcalls = nf;
    return res;
}

/* find_2pwl_hardening_norm -> _get_hardening_norm_2pwl (sam_cyutils.pyx:289-353) */
void or_find_2pwl_norm(double target_time, const double *mtot, const double *mrat, int64_t n,
                       double sepa_init, double rchar, double gamma_inner, double gamma_outer, int nsteps,
                       double *norm_log10) {
    const double XTOL = 1e-3, RTOL = 1e-5, LO = -20.0, HI = +20.0;
    const int MITR = 100;
    for (int64_t ii = 0; ii < n; ii++) {
        norm_log10[ii] = or_brentq_lifetime(LO, HI, XTOL, RTOL, MITR, mtot[ii], mrat[ii], target_time, sepa_init,
                                            rchar, gamma_inner, gamma_outer, nsteps, NULL, NULL);
    }
}

/* ---------------------------------------------------------------------------
 * _dynamic_binary_number_at_fobs_2pwl  (sam_cyutils.pyx:506-781)
 * hard_norm (M,Q), nden (M,Q,Z), gmt_time (M,Q,Z); outputs (M,Q,Z,F) must be
 * pre-initialised by the caller to redz_final=-1, diff_num=0 (sam_cyutils.pyx:460-461).
 * ------------------------------------------------------------------------- */
int or_dbn_2pwl(const double *target_fobs_orb, int n_freq, double sepa_init, int num_steps,
                const double *hard_norm, double hard_rchar, double hard_gamma_inner, double hard_gamma_outer,
                const double *nden, const double *mtot, int n_mtot, const double *mrat, int n_mrat,
                const double *redz, int n_redz, const double *gmt_time,
                const double *redz_interp_grid, const double *dcom_interp_grid, const double *tage_interp_grid,
                int n_interp, double *redz_final, double *diff_num) {
    or_init();
    double age_universe = tage_interp_grid[n_interp - 1];
    double sepa_init_log10 = log10(sepa_init);
    double *redz_age = (double *)malloc(sizeof(double) * n_redz);
    if (!redz_age) return -1;

    int ii = 0;
    for (int kk = 0; kk < n_redz; kk++) {
        int rev = n_redz - 1 - kk;
        while ((ii + 1 < n_interp) && (redz_interp_grid[ii + 1] > redz[rev]) && (ii < n_interp - 1)) ii += 1;
        redz_age[rev] = interp_at_index(ii, redz[rev], redz_interp_grid, tage_interp_grid);
    }

    for (ii = 0; ii < n_mtot; ii++) {
        double mt = mtot[ii];
        double risco = 3.0 * MY_SCHW * mt;
        double dx = (sepa_init_log10 - log10(risco)) / num_steps;
        for (int jj = 0; jj < n_mrat; jj++) {
            double mr = mrat[jj];
            double sepa_log10 = sepa_init_log10;
            double norm = hard_norm[(size_t)ii * n_mrat + jj];
            double sepa_left = pow(10.0, sepa_log10);
            double dadt_left = or_hard_func_2pwl_gw(mt, mr, sepa_left, norm, hard_rchar, hard_gamma_inner, hard_gamma_outer);
            double frst_orb_left = kepler_freq_from_sepa(mt, sepa_left);
            double time_evo = 0.0;
            int interp_left_idx = 0;
            for (int step = 0; step < num_steps; step++) {
                sepa_log10 -= dx;
                double sepa_right = pow(10.0, sepa_log10);
                double frst_orb_right = kepler_freq_from_sepa(mt, sepa_right);
                double dadt_right = or_hard_func_2pwl_gw(mt, mr, sepa_right, norm, hard_rchar, hard_gamma_inner, hard_gamma_outer);
                double dt = 2.0 * (sepa_right - sepa_left) / (dadt_left + dadt_right);
                time_evo += dt;
                for (int kk = n_redz - 1; kk > -1; kk--) {
                    size_t cell = ((size_t)ii * n_mrat + jj) * n_redz + kk;
                    double gmt = gmt_time[cell];
                    double time_right = time_evo + gmt + redz_age[kk];
                    double time_left = time_right - dt;
                    if (time_left > age_universe) continue;
                    interp_left_idx = ww_increasing(interp_left_idx, n_interp, time_left, tage_interp_grid);
                    double redz_left = interp_at_index(interp_left_idx, time_left, tage_interp_grid, redz_interp_grid);
                    if (redz_left < 0.0) continue;
                    int interp_right_idx = ww_increasing(interp_left_idx, n_interp, time_right, tage_interp_grid);
                    double redz_right = interp_at_index(interp_right_idx, time_right, tage_interp_grid, redz_interp_grid);
                    if (redz_right < 0.0) redz_right = 0.0;
                    double fobs_orb_left = frst_orb_left / (1.0 + redz_left);
                    double fobs_orb_right = frst_orb_right / (1.0 + redz_right);
                    for (int ff = 0; ff < n_freq; ff++) {
                        double ftarget = target_fobs_orb[ff];
                        if ((ftarget < fobs_orb_left) || (fobs_orb_right < ftarget)) continue;
                        double new_time = interp_between(ftarget, fobs_orb_left, fobs_orb_right, time_left, time_right);
                        if (new_time > tage_interp_grid[n_interp - 1]) break;
                        int new_interp_idx = ww_increasing(interp_left_idx, n_interp, new_time, tage_interp_grid);
                        double new_redz = interp_at_index(new_interp_idx, new_time, tage_interp_grid, redz_interp_grid);
                        double dcom = interp_at_index(new_interp_idx, new_time, tage_interp_grid, dcom_interp_grid);
                        redz_final[cell * n_freq + ff] = new_redz;
                        double target_frst_orb = ftarget * (1.0 + new_redz);
                        double sepa = kepler_sepa_from_freq(mt, target_frst_orb);
                        double dadt = or_hard_func_2pwl_gw(mt, mr, sepa, norm, hard_rchar, hard_gamma_inner, hard_gamma_outer);
                        double tres = -(2.0 / 3.0) * sepa / dadt;
                        double cosmo_fact = FOUR_PI_SPLC_OVER_MPC * (1.0 + new_redz) * pow(dcom / MY_MPC, 2);
                        diff_num[cell * n_freq + ff] = nden[cell] * tres * cosmo_fact;
                    }
                }
                dadt_left = dadt_right;
                sepa_left = sepa_right;
                frst_orb_left = frst_orb_right;
            }
        }
    }
    free(redz_age);
    return 0;
}

/* ---------------------------------------------------------------------------
 * _dynamic_binary_number_at_fobs_gw  (sam_cyutils.pyx:784-899)
 * ------------------------------------------------------------------------- */
int or_dbn_gw(const double *target_fobs_orb, int n_freq, const double *nden, const double *mtot, int n_mtot,
              const double *mrat, int n_mrat, int n_redz, const double *redz_prime,
              const double *redz_interp_grid, const double *dcom_interp_grid, int n_interp,
              double *redz_final, double *diff_num) {
    or_init();
    for (int ii = 0; ii < n_mtot; ii++) {
        double mt = mtot[ii];
        double rad_isco = 3.0 * MY_SCHW * mt;
        double frst_orb_isco = kepler_freq_from_sepa(mt, rad_isco);
        for (int jj = 0; jj < n_mrat; jj++) {
            double mr = mrat[jj];
            int interp_idx = 0;
            for (int _kk = 0; _kk < n_redz; _kk++) {
                int kk = n_redz - 1 - _kk;
                size_t cell = ((size_t)ii * n_mrat + jj) * n_redz + kk;
                double rzp = redz_prime[cell];
                if (rzp <= 0.0) continue;
                for (int ff = 0; ff < n_freq; ff++) {
                    redz_final[cell * n_freq + ff] = rzp;
                    double ftarget = target_fobs_orb[ff];
                    double target_frst_orb = ftarget * (1.0 + rzp);
                    if (target_frst_orb > frst_orb_isco) {
                        for (int fp = ff + 1; fp < n_freq; fp++) redz_final[cell * n_freq + fp] = rzp;
                        break;
                    }
                    interp_idx = ww_decreasing(interp_idx, n_interp, rzp, redz_interp_grid);
                    double dcom = interp_at_index(interp_idx, rzp, redz_interp_grid, dcom_interp_grid);
                    double sepa = kepler_sepa_from_freq(mt, target_frst_orb);
                    double dadt = or_hard_gw(mt, mr, sepa);
                    double tres = -(2.0 / 3.0) * sepa / dadt;
                    double cosmo_fact = FOUR_PI_SPLC_OVER_MPC * (1.0 + rzp) * pow(dcom / MY_MPC, 2);
                    diff_num[cell * n_freq + ff] = nden[cell] * tres * cosmo_fact;
                }
            }
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------------
 * Realization walks with INJECTED counts.
 * counts: (B, F, R) C-order doubles (a Poisson or Normal draw per bin,
 * frequency, realization, exactly what the reference draws inline);
 * h2fdf: (B, F).  Bins are flat C-order indices over (M-1, Q-1, Z-1).
 * ------------------------------------------------------------------------- */

/*
 * _loudest_hc_from_sorted (cyutils.pyx:1266-1344) and, with skip_zero_h=1 and
 * par != NULL, _loudest_hc_and_par_from_sorted_redz (cyutils.pyx:1615-1767).
 * order: (B,) flat bin indices, loudest first.
 * Outputs: hc2ss (F,R,L) [zero-filled by caller], hc2bg (F,R), ssidx (F,R,L)
 * flat bin index or -1 [filled with -1 by caller], bgpar (npar,F,R) (may be NULL):
 * strain-weighted means  sum(n*h*par)/sum(n*h) of the background, par (npar,B,F).
 */
void or_loudest_counts(const double *counts, const double *h2fdf, const int64_t *order, int64_t B, int F, int R,
                       int L, int skip_zero_h, const double *par, int npar, double *hc2ss, double *hc2bg,
                       int64_t *ssidx, double *bgpar) {
    double acc[16];
    for (int rr = 0; rr < R; rr++) {
        for (int ff = 0; ff < F; ff++) {
            int ll = 0;
            double sum = 0.0;
            for (int k = 0; k < npar; k++) acc[k] = 0.0;
            for (int64_t bb = 0; bb < B; bb++) {
                int64_t b = order[bb];
                double num = counts[((size_t)b * F + ff) * R + rr];
                double cur = h2fdf[(size_t)b * F + ff];
                if (skip_zero_h) {
                    if ((num < 1) || (cur == 0)) continue;
                } else {
                    if (num < 1) continue;
                }
                while ((ll < L) && (num > 0)) {
                    size_t o = ((size_t)ff * R + rr) * L + ll;
                    hc2ss[o] = cur;
                    ssidx[o] = b;
                    num -= 1;
                    ll += 1;
                }
                sum += num * cur;
                for (int k = 0; k < npar; k++) acc[k] += num * cur * par[((size_t)k * B + b) * F + ff];
            }
            hc2bg[(size_t)ff * R + rr] = sum;
            if (bgpar)
                for (int k = 0; k < npar; k++) bgpar[((size_t)k * F + ff) * R + rr] = acc[k] / sum;
        }
    }
}

/*
 * _ss_bg_hc (cyutils.pyx:935-1014) and _ss_bg_hc_and_par (cyutils.pyx:1068-1178):
 * per-frequency argmax over bins in C order with strict '>' among bins with n>0;
 * hc2bg = sum(n*h) - max.  ssidx (F,R) flat bin index (-1 if none).
 * bgpar (npar,F,R) = (sum(n*h*p) - h_max*p_max)/(sum - max), par (npar, B).
 * Where the reference raises (no source found, cyutils.pyx:1157-1158) this
 * restatement stores ssidx=-1 and bgpar=NaN and continues.
 */
void or_ss_bg_counts(const double *counts, const double *h2fdf, int64_t B, int F, int R, const double *par, int npar,
                     double *hc2ss, double *hc2bg, int64_t *ssidx, double *bgpar) {
    double acc[16];
    for (int rr = 0; rr < R; rr++) {
        for (int ff = 0; ff < F; ff++) {
            double mx = 0.0, sum = 0.0;
            int64_t imax = -1;
            for (int k = 0; k < npar; k++) acc[k] = 0.0;
            for (int64_t b = 0; b < B; b++) {
                double num = counts[((size_t)b * F + ff) * R + rr];
                double cur = h2fdf[(size_t)b * F + ff];
                if (cur > mx && num > 0) {
                    mx = cur;
                    imax = b;
                }
                sum += num * cur;
                for (int k = 0; k < npar; k++) acc[k] += num * cur * par[(size_t)k * B + b];
            }
            hc2ss[(size_t)ff * R + rr] = mx;
            hc2bg[(size_t)ff * R + rr] = sum - mx;
            ssidx[(size_t)ff * R + rr] = imax;
            if (bgpar) {
                for (int k = 0; k < npar; k++) {
                    double v = (imax < 0) ? NAN : (acc[k] - h2fdf[(size_t)imax * F + ff] * par[(size_t)k * B + imax]) / (sum - mx);
                    bgpar[((size_t)k * F + ff) * R + rr] = v;
                }
            }
        }
    }
}

/* _sam_poisson_gwb (cyutils.pyx:862-897) with injected counts: gwb (F,R) */
void or_poisson_gwb_counts(const double *counts, const double *h2fdf, int64_t B, int F, int R, double *gwb) {
    for (size_t i = 0; i < (size_t)F * R; i++) gwb[i] = 0.0;
    for (int64_t b = 0; b < B; b++)
        for (int ff = 0; ff < F; ff++)
            for (int rr = 0; rr < R; rr++)
                gwb[(size_t)ff * R + rr] += counts[((size_t)b * F + ff) * R + rr] * h2fdf[(size_t)b * F + ff];
}

/* ---------------------------------------------------------------------------
 * Random realizations for the CPU baseline: the reference loop structure
 * (one draw per bin, frequency, realization; Normal above the threshold),
 * with xoshiro256** (Blackman & Vigna 2018) in place of numpy's PCG64 and the
 * Poisson samplers numpy documents: multiplication (Knuth) below 10 and PTRS
 * (Hoermann 1993) at and above 10.
 * ------------------------------------------------------------------------- */
typedef struct { uint64_t s[4]; } xo_state;

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t xo_next(xo_state *st) {
    uint64_t *s = st->s;
    uint64_t result = rotl64(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}
static double xo_double(xo_state *st) { return (xo_next(st) >> 11) * (1.0 / 9007199254740992.0); }
static void xo_seed(xo_state *st, uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; i++) {
        z += 0x9e3779b97f4a7c15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        st->s[i] = x ^ (x >> 31);
    }
}
static double xo_gauss(xo_state *st) {
    double x1, x2, r2;
    do {
        x1 = 2.0 * xo_double(st) - 1.0;
        x2 = 2.0 * xo_double(st) - 1.0;
        r2 = x1 * x1 + x2 * x2;
    } while (r2 >= 1.0 || r2 == 0.0);
    return x2 * sqrt(-2.0 * log(r2) / r2);
}
static int64_t poisson_mult(xo_state *st, double lam) {
    double enlam = exp(-lam), prod = 1.0;
    int64_t X = 0;
    for (;;) {
        prod *= xo_double(st);
        if (prod > enlam) X += 1;
        else return X;
    }
}
static int64_t poisson_ptrs(xo_state *st, double lam) {
    double slam = sqrt(lam), loglam = log(lam);
    double b = 0.931 + 2.53 * slam;
    double a = -0.059 + 0.02483 * b;
    double invalpha = 1.1239 + 1.1328 / (b - 3.4);
    double vr = 0.9277 - 3.6224 / (b - 2);
    for (;;) {
        double U = xo_double(st) - 0.5;
        double V = xo_double(st);
        double us = 0.5 - fabs(U);
        int64_t k = (int64_t)floor((2 * a / us + b) * U + lam + 0.43);
        if ((us >= 0.07) && (V <= vr)) return k;
        if ((k < 0) || ((us < 0.013) && (V > us))) continue;
        if ((log(V) + log(invalpha) - log(a / (us * us) + b)) <= (-lam + k * loglam - lgamma((double)k + 1))) return k;
    }
}
static double draw_count(xo_state *st, double lam, double thresh) {
    if (lam > thresh) return lam + sqrt(lam) * xo_gauss(st);
    if (lam >= 10) return (double)poisson_ptrs(st, lam);
    if (lam == 0) return 0.0;
    return (double)poisson_mult(st, lam);
}

/* Draw counts (B,F,R) with the sampler above (CPU statistical oracle). */
void or_draw_counts(const double *number, int64_t B, int F, int R, double thresh, uint64_t seed, double *counts) {
    xo_state st;
    xo_seed(&st, seed);
    for (int64_t b = 0; b < B; b++)
        for (int ff = 0; ff < F; ff++) {
            double lam = number[(size_t)b * F + ff];
            for (int rr = 0; rr < R; rr++) counts[((size_t)b * F + ff) * R + rr] = draw_count(&st, lam, thresh);
        }
}

/*
 * The reference's realization loop (cyutils.pyx:1318-1344) with inline draws:
 * the timed CPU baseline.  Same outputs as or_loudest_counts.
 */
void or_loudest_rng(const double *number, const double *h2fdf, const int64_t *order, int64_t B, int F, int R,
                    int L, double thresh, uint64_t seed, double *hc2ss, double *hc2bg) {
    xo_state st;
    xo_seed(&st, seed);
    for (int rr = 0; rr < R; rr++) {
        for (int ff = 0; ff < F; ff++) {
            int ll = 0;
            double sum = 0.0;
            for (int64_t bb = 0; bb < B; bb++) {
                int64_t b = order[bb];
                double num = draw_count(&st, number[(size_t)b * F + ff], thresh);
                if (num < 1) continue;
                double cur = h2fdf[(size_t)b * F + ff];
                while ((ll < L) && (num > 0)) {
                    hc2ss[((size_t)ff * R + rr) * L + ll] = cur;
                    num -= 1;
                    ll += 1;
                }
                sum += num * cur;
            }
            hc2bg[(size_t)ff * R + rr] = sum;
        }
    }
}

/* Residence time f/(df/dt) = -(2/3) a / (da/dt) at rest-frame orbital
 * frequency frst under GW hardening, as in the inner loop of
 * _dynamic_binary_number_at_fobs_gw (sam_cyutils.pyx:886-891). */
double or_gw_tres(double mt, double mr, double frst) {
    or_init();
    double sepa = kepler_sepa_from_freq(mt, frst);
    double dadt = or_hard_gw(mt, mr, sepa);
    return -(2.0 / 3.0) * sepa / dadt;
}

double or_kepler_sepa(double mt, double frst) {
    or_init();
    return kepler_sepa_from_freq(mt, frst);
}
