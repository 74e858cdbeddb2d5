"""Oracle of the eccentric SAM GWB: the eccentricity evolution of sams/sam.py:1235-1288 and the
harmonic sums of cyutils.pyx:361-851 (``sam_calc_gwb_single_eccen`` and its ``_discrete``
Poisson variant), restated in numpy with scipy.special.jv for the two Bessel values per
(harmonic, eccentricity), as the reference calls it (cyutils.pyx:181-182).

The loops follow the reference's order and control flow: redshift, total mass in reverse,
mass ratios (vectorised here), then the (frequency, harmonic) pairs in ascending observer
frequency with the same evolution-index walk, ``continue`` below the first evolution point and
``break`` above the last.  Pure-Python loops: small grids only.

TEST INFRASTRUCTURE ONLY.
"""
import numpy as np
import scipy.special as sps

# cyutils.pyx:39-48 (the Cython module's own constants)
ECCEN_ZERO_LIMIT = 1.0e-12
MY_NWTG = 6.6742999e-08
MY_SPLC = 29979245800.0
MY_MPC = 3.08567758e+24
GW_DADT_SEP_CONST = - 64.0 * pow(MY_NWTG, 3) / 5.0 / pow(MY_SPLC, 5)
GW_SRC_CONST = 8.0 * pow(MY_NWTG, 5.0/3.0) * pow(np.pi, 2.0/3.0) / np.sqrt(10.0) / pow(MY_SPLC, 4.0)


def gw_dade(sepa, eccen):
    """utils.gw_dade (utils.py:2077-2102), Peters (1964) Eq. 5.7."""
    e2 = eccen ** 2
    num = (1 + (73.0 / 24.0) * e2 + (37.0 / 96.0) * e2 * e2)
    den = (1 - e2) * (1.0 + (121.0 / 304.0) * e2)
    with np.errstate(divide='ignore', invalid='ignore'):
        return (12.0 / 19.0) * (sepa / eccen) * (num / den)


def deda(sepa, eccen):
    """hardening.Hard_GW.deda (hardening.py:185-208): 1 / (da/de)."""
    with np.errstate(divide='ignore'):
        return 1.0 / gw_dade(sepa, eccen)


def evolve_eccen_uniform_single(mtot, eccen_init, sepa_init, nsteps, schw):
    """sams/sam.py:1235-1288: RK4 steps (utils.rk4_step, utils.py:1018-1043) of de/da in
    log-spaced separation from ``sepa_init`` down to 3 Schwarzschild radii of the smallest
    total mass; ``schw`` = 2G/c^2 of the reference's constants."""
    assert (0.0 <= eccen_init) and (eccen_init < 1.0)
    eccen = np.zeros(nsteps)
    eccen[0] = eccen_init
    sepa_coal = schw * np.asarray(mtot) * 3
    sepa_min = sepa_coal.min()
    sepa = np.logspace(*np.log10([sepa_init, sepa_min]), nsteps)
    for step in range(1, nsteps):
        a0 = sepa[step - 1]
        a1 = sepa[step]
        da = (a1 - a0)
        e0 = eccen[step - 1]
        k1 = da * deda(a0, e0)
        k2 = da * deda(a0 + da / 2.0, e0 + k1 / 2.0)
        k3 = da * deda(a0 + da / 2.0, e0 + k2 / 2.0)
        k4 = da * deda(a0 + da, e0 + k3)
        e1 = e0 + (1.0 / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        eccen[step] = np.clip(e1, 0.0, None)
    return sepa, eccen


def gw_freq_dist_func(nn, ee):
    """g(n, e), cyutils.pyx:145-199: J_{n-2}, J_{n-1} from scipy, J_n..J_{n+2} by the
    upward recurrence (cyutils.pyx:54-76); e below 1e-12 is circular."""
    if ee < ECCEN_ZERO_LIMIT:
        return 1.0 if nn == 2 else 0.0
    ne = nn * ee
    n2 = nn * nn
    jn_m2 = sps.jv(nn - 2, ne)
    jn_m1 = sps.jv(nn - 1, ne)
    jn = (2 * (nn - 1) / ne) * jn_m1 - jn_m2
    jn_p1 = (2 * (nn + 1 - 1) / ne) * jn - jn_m1
    jn_p2 = (2 * (nn + 2 - 1) / ne) * jn_p1 - jn
    aa = jn_m2 - 2.0 * ee * jn_m1 + (2 / nn) * jn + 2 * ee * jn_p1 - jn_p2
    aa = aa * aa
    bb = jn_m2 - 2 * ee * jn + jn_p2
    bb = (1 - ee * ee) * bb * bb
    cc = (4.0 / (3.0 * n2)) * jn * jn
    return (n2 * n2 / 32) * (aa + bb + cc)


def gw_ecc_func(eccen):
    """F(e), cyutils.pyx:79-99 (Peters 1964 Eq. 5.6)."""
    e2 = eccen * eccen
    return (1.0 + (73.0 / 24.0) * e2 + (37.0 / 96.0) * e2 * e2) / pow(1.0 - e2, 7.0 / 2.0)


def trapz_grid_weight(index, size, grid):
    """cyutils.pyx:102-142: (inverse weight, width) of a grid point."""
    if index == 0:
        return 2.0, grid[1] - grid[0]
    if index == size - 1:
        return 2.0, grid[index] - grid[index - 1]
    return 1.0, 0.5 * (grid[index + 1] - grid[index - 1])


def _walk(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms, visit):
    """The loop nest shared by both reference kernels (cyutils.pyx:466-585, 713-838); calls
    visit(ii, kk, aa, bb, nh, gwfr, sa_fourth, ecc, mt, m1, m2, mchirp, zterm, dc_cm, dc_term, jw)
    with jw the mass-ratio weights (Q, 2) for every contributing (mass, redshift, f, harmonic).

    (mass, redshift) cells whose density is zero at every mass ratio are not walked: every term
    they would add is proportional to ndens[ii, :, kk], so they add exact zeros (and Poisson means
    of zero).  The evolution index a skipped cell would have recorded at its first (f, harmonic)
    is only the starting point of the next, lighter mass's forward walk; the index from a
    heavier mass is a lower bound of it (rest-frame frequencies of a lighter binary reach each
    separation later), and the forward walk from any lower bound stops at the same index."""
    nfreqs, n_mtot, n_mrat, n_redz, n_eccs = len(gwfobs), len(mtot_log10), len(mrat), len(redz), len(sepa_evo)
    kep_sa_term = MY_NWTG / pow(2.0 * np.pi, 2)
    freq_pref = (1.0 / (2.0 * np.pi)) * np.sqrt(MY_NWTG)
    frst_evo_pref = freq_pref / np.power(np.asarray(sepa_evo, float), 1.5)
    bb_, aa_ = np.meshgrid(np.arange(nharms), np.arange(nfreqs))
    freq_harms = (np.asarray(gwfobs)[:, None] / (np.arange(nharms)[None, :] + 1)).ravel()
    order = np.argsort(freq_harms, kind='stable')
    jw = np.array([trapz_grid_weight(jj, n_mrat, mrat) for jj in range(n_mrat)])
    mrat = np.asarray(mrat, float)
    for kk in range(n_redz):
        zterm = 1.0 + redz[kk]
        dc_mpc = dcom[kk]
        dc_cm = dc_mpc * MY_MPC
        dc_term = 4 * np.pi * (MY_SPLC / MY_MPC) * pow(dc_mpc, 2)
        ecc_idx_beg = 0
        for ii_mm in range(n_mtot):
            ii = n_mtot - 1 - ii_mm
            if not np.any(ndens[ii, :, kk]):
                continue
            mt = pow(10.0, mtot_log10[ii])
            mt_sqrt = np.sqrt(mt)
            kep_sa_mass_term = kep_sa_term * mt
            m1 = mt / (1.0 + mrat)
            m2 = mt - m1
            mchirp = mt * np.power(mrat, 3.0 / 5.0) / np.power(1 + mrat, 6.0 / 5.0)
            ecc_idx = ecc_idx_beg
            for ff in range(nfreqs * nharms):
                aa, bb = divmod(int(order[ff]), nharms)
                nh = bb + 1
                gwfr = gwfobs[aa] * zterm / nh
                sa = pow(kep_sa_mass_term / pow(gwfr, 2), 1.0 / 3.0)
                sa_fourth = pow(sa, 4)
                frst_evo_lo = frst_evo_pref[ecc_idx] * mt_sqrt
                frst_evo_hi = frst_evo_pref[ecc_idx + 1] * mt_sqrt
                while (gwfr > frst_evo_hi) and (ecc_idx < n_eccs - 2):
                    frst_evo_lo = frst_evo_hi
                    ecc_idx += 1
                    frst_evo_hi = frst_evo_pref[ecc_idx + 1] * mt_sqrt
                if ff == 0:
                    ecc_idx_beg = ecc_idx
                if gwfr < frst_evo_lo:
                    continue
                if gwfr > frst_evo_hi:
                    break
                ecc = (eccen_evo[ecc_idx + 1] - eccen_evo[ecc_idx]) / (frst_evo_hi - frst_evo_lo)
                ecc = eccen_evo[ecc_idx] + (gwfr - frst_evo_lo) * ecc
                visit(ii, kk, aa, bb, nh, gwfr, sa_fourth, ecc, mt, m1, m2, mchirp, zterm, dc_cm, dc_term, jw)


def _ik_weights(ii, kk, mtot_log10, redz):
    return trapz_grid_weight(ii, len(mtot_log10), mtot_log10), trapz_grid_weight(kk, len(redz), redz)


def sam_calc_gwb_single_eccen(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms=100):
    """cyutils.pyx:361-597: (F, H) characteristic strain squared per frequency and harmonic."""
    ndens = np.asarray(ndens, float)
    gwb = np.zeros((len(gwfobs), nharms))

    def visit(ii, kk, aa, bb, nh, gwfr, sa_fourth, ecc, mt, m1, m2, mchirp, zterm, dc_cm, dc_term, jw):
        ival, kval = _ik_weights(ii, kk, mtot_log10, redz)
        weight_ik = ival[1] * kval[1] / (ival[0] * kval[0])
        weight = weight_ik * (jw[:, 1] / jw[:, 0])
        hterm_pref = ndens[ii, :, kk] * dc_term * zterm
        hterm_pref = hterm_pref * np.power(GW_SRC_CONST * mchirp * np.power(2.0 * mchirp, 2.0 / 3.0) / dc_cm, 2)
        gne = gw_freq_dist_func(nh, ecc)
        fe_ecc = gw_ecc_func(ecc)
        tau = - sa_fourth / (GW_DADT_SEP_CONST * fe_ecc * m1 * m2 * mt)
        hterm = hterm_pref * tau * gne * (4.0 / (nh * nh)) * pow(gwfr, 4.0 / 3.0)
        gwb[aa, bb] += np.sum(hterm * weight)

    _walk(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms, visit)
    return gwb


def eccen_number_h(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms=100):
    """Per-(node, frequency, harmonic) Poisson mean and strain factor of the discrete variant
    (cyutils.pyx:766-838): ``number_term`` and ``hterm / weight`` as (M, Q, Z, F, H) arrays, zero
    where the reference draws nothing or adds only zeros (outside the evolution, g(n,e) = 0)."""
    ndens = np.asarray(ndens, float)
    M, Q, Z = ndens.shape
    lam = np.zeros((M, Q, Z, len(gwfobs), nharms))
    hval = np.zeros_like(lam)

    def visit(ii, kk, aa, bb, nh, gwfr, sa_fourth, ecc, mt, m1, m2, mchirp, zterm, dc_cm, dc_term, jw):
        ival, kval = _ik_weights(ii, kk, mtot_log10, redz)
        volume = ival[1] * kval[1] * jw[:, 1]
        weight = ival[0] * kval[0] * jw[:, 0]
        number_term_pref = ndens[ii, :, kk] * dc_term * zterm
        hterm_pref = np.power(GW_SRC_CONST * mchirp * np.power(2.0 * mchirp, 2.0 / 3.0) / dc_cm, 2)
        gne = gw_freq_dist_func(nh, ecc)
        fe_ecc = gw_ecc_func(ecc)
        tau = - sa_fourth / (GW_DADT_SEP_CONST * fe_ecc * m1 * m2 * mt)
        number_term = number_term_pref * tau * volume
        hterm = hterm_pref * gne * (4.0 / (nh * nh)) * pow(gwfr, 4.0 / 3.0)
        h = hterm / weight
        lam[ii, :, kk, aa, bb] = np.where(h != 0.0, number_term, 0.0)
        hval[ii, :, kk, aa, bb] = h

    _walk(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms, visit)
    return lam, hval


def sam_calc_gwb_single_eccen_discrete(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms,
                                       nreals, seed=0):
    """cyutils.pyx:600-851 with numpy's Poisson draws: (F, H, R)."""
    lam, hval = eccen_number_h(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms)
    rng = np.random.default_rng(seed)
    F, H = lam.shape[3:]
    lam2 = lam.reshape(-1, F, H)
    h2 = hval.reshape(-1, F, H)
    out = np.zeros((F, H, nreals))
    for rr in range(nreals):
        out[:, :, rr] = np.sum(h2 * rng.poisson(lam2), axis=0)
    return out
