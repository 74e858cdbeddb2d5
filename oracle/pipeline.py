"""Oracle pipeline: numpy restatements + ctypes to liboracle.so (FP64 C restatement).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Each function cites the reference lines it restates.  What pins each part is
listed in tests/ (golden vectors of holodeck/tests/test_utils.py:60-94,
scipy.optimize.brentq, mpmath closed forms, brute force on small grids).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

# astropy constants used by the reference's Python modules (holodeck/constants.py:19-47)
NWTG = 6.6743e-08
SPLC = 29979245800.0
MSOL = 1.988409870698051e+33
PC = 3.0856775814913673e+18
MPC = 1.0e6 * PC
YR = 31557600.0
GYR = 1.0e9 * YR
_GW_SRC_CONST = 8 * np.power(NWTG, 5 / 3) * np.power(np.pi, 2 / 3) / np.sqrt(10) / np.power(SPLC, 4)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        d, i, i64, p, u64 = C.c_double, C.c_int, C.c_int64, C.c_void_p, C.c_uint64
        sigs = {
            "or_constants": (None, [p]),
            "or_hard_gw": (d, [d, d, d]),
            "or_hard_func_2pwl_gw": (d, [d, d, d, d, d, d, d]),
            "or_lifetime_2pwl": (d, [d, d, d, d, d, d, d, d, i]),
            "or_brentq_lifetime": (d, [d, d, d, d, i, d, d, d, d, d, d, d, i, p, p]),
            "or_find_2pwl_norm": (None, [d, p, p, i64, d, d, d, d, i, p]),
            "or_integrate_3dx1d": (None, [p, i, p, i, p, i, p, i, p, p]),
            "or_dbn_2pwl": (i, [p, i, d, i, p, d, d, d, p, p, i, p, i, p, i, p, p, p, p, i, p, p]),
            "or_dbn_gw": (i, [p, i, p, p, i, p, i, i, p, p, p, i, p, p]),
            "or_loudest_counts": (None, [p, p, p, i64, i, i, i, i, p, i, p, p, p, p]),
            "or_ss_bg_counts": (None, [p, p, i64, i, i, p, i, p, p, p, p]),
            "or_poisson_gwb_counts": (None, [p, p, i64, i, i, p]),
            "or_draw_counts": (None, [p, i64, i, i, d, u64, p]),
            "or_loudest_rng": (None, [p, p, p, i64, i, i, i, d, u64, p, p]),
            "or_gw_tres": (d, [d, d, d]),
            "or_kepler_sepa": (d, [d, d]),
        }
        for k, (r, a) in sigs.items():
            fn = getattr(L, k)
            fn.restype = r
            fn.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------------------
# Cosmology: flat LambdaCDM without radiation (astropy FlatLambdaCDM, Tcmb0=0)
# ---------------------------------------------------------------------------
class OracleCosmo:
    def __init__(self, h=0.6933, Om0=0.2880, size=200, zmax=20.0):
        self.h = h
        self.Om0 = Om0
        self.Ode0 = 1.0 - Om0
        self.H0 = h * 100.0 * 1.0e5 / MPC
        self.t_H = 1.0 / self.H0
        self.d_H = SPLC / self.H0
        xx = np.linspace(np.log10(1.0 + zmax), 0.0, size)
        zz = 10.0 ** xx - 1.0
        zz[-1] = 0.0
        self._grid_z = zz
        self._grid_age = self.age(zz)
        self._grid_dcom = self.comoving_distance(zz)
        self.size = size

    def age(self, z):
        z = np.asarray(z, dtype=float)
        return self.t_H * (2.0 / 3.0) / np.sqrt(self.Ode0) * np.arcsinh(
            np.sqrt(self.Ode0 / self.Om0) * (1.0 + z) ** -1.5)

    def tage_to_z(self, t):
        s = np.sinh(1.5 * np.sqrt(self.Ode0) * np.asarray(t, dtype=float) / self.t_H)
        return (np.sqrt(self.Ode0 / self.Om0) / s) ** (2.0 / 3.0) - 1.0

    def dtdz(self, z):
        z = np.asarray(z, dtype=float)
        return self.t_H / ((1.0 + z) * np.sqrt(self.Om0 * (1.0 + z) ** 3 + self.Ode0))

    def comoving_distance(self, z):
        """d_H * int_0^z dz'/E(z') via u=(1+z)^-1/2: int_u^1 2du/sqrt(Om + Ode u^6), 64-pt Gauss-Legendre."""
        z = np.asarray(z, dtype=float)
        x, w = np.polynomial.legendre.leggauss(64)
        u0 = 1.0 / np.sqrt(1.0 + z)
        c = 0.5 * (1.0 - u0)
        m = 0.5 * (1.0 + u0)
        uu = m[..., None] + c[..., None] * x
        f = 2.0 / np.sqrt(self.Om0 + self.Ode0 * uu ** 6)
        return self.d_H * c * np.sum(w * f, axis=-1)

    @property
    def age_now(self):
        return float(self.age(0.0))


# ---------------------------------------------------------------------------
# Static binary density (sam.py:250-398) without scatter
# ---------------------------------------------------------------------------
def mstar_from_mbh(mbh, mmb):
    """MMBulge_Standard.mstar_from_mbh (host_relations.py:768-771, 1137-1173; BF_Constant 190-192)."""
    xx = np.log10(mbh / mmb['mamp'])
    xx = (1.0 / mmb['mplaw']) * xx
    mbulge = mmb['mref'] * np.power(10.0, xx)
    return mbulge / mmb['bulge_frac']


def static_density(mtot, mrat, redz, cfg, cosmo):
    """Returns (dens [before scatter, before stalled zeroing], zprime or None, gmt_time or None)."""
    mmb = cfg['mmbulge']
    rz = redz[np.newaxis, np.newaxis, :]
    m1 = mtot[:, None] / (1.0 + mrat[None, :])
    m2 = mtot[:, None] - m1
    mbh_pri, mbh_sec, rz = np.broadcast_arrays(m1[..., None], m2[..., None], rz)
    mstar_pri = mstar_from_mbh(mbh_pri, mmb)
    mstar_sec = mstar_from_mbh(mbh_sec, mmb)
    mstar_rat = mstar_sec / mstar_pri
    mstar_tot = mstar_pri + mstar_sec

    zprime = gmt_time = None
    if cfg.get('gmt') is not None:
        g = cfg['gmt']
        tau0 = g['time_norm'] * np.power(mstar_pri / g['mref'], g['malpha']) * np.power(1.0 + rz, g['zbeta']) \
            * np.power(mstar_rat, g['qgamma'])
        # utils.redz_after (utils.py:1772-1808)
        new_age = cosmo.age(rz) + tau0
        zprime = -1.0 * np.ones_like(new_age)
        idx = new_age < cosmo.age_now
        zprime[idx] = cosmo.tage_to_z(new_age[idx])
        gmt_time = tau0

    if cfg.get('gmr') is None:
        g = cfg['gpf']
        rv = g['frac_norm'] * np.power(mstar_pri / g['mref'], g['malpha']) * np.power(1.0 + rz, g['zbeta']) \
            * np.power(mstar_rat, g['qgamma'])
        rv = np.clip(rv, None, g['max_frac'])
        rate = rv / gmt_time
    else:
        g = cfg['gmr']
        zp1 = 1.0 + rz
        norm = g['norm0'] * np.power(zp1, g['normz'])
        malpha = g['malpha0'] * np.power(zp1, g['malphaz'])
        mdelta = g['mdelta0'] * np.power(zp1, g['mdeltaz'])
        qgamma = g['qgamma0'] * np.power(zp1, g['qgammaz']) + g['qgammam'] * np.log10(mstar_tot / g['mref'])
        rate = norm * np.power(mstar_tot / g['mref'], malpha) * np.power(1.0 + mstar_tot / g['mref_delta'], mdelta) \
            * np.power(mstar_rat, qgamma)

    s = cfg['gsmf']
    phi = np.power(10.0, s['phi0'] + s['phiz'] * rz)
    mchar = s['mchar0'] + s['mcharz'] * rz
    alpha = s['alpha0'] + s['alphaz'] * rz
    xx = mstar_pri / mchar
    gsmf = np.log(10.0) * phi * np.power(xx, 1.0 + alpha) * np.exp(-xx)
    dens = gsmf * rate * cosmo.dtdz(rz)

    plaw = mmb['mplaw']
    dqbh_dqgal = plaw * np.power(mstar_rat, plaw - 1.0)
    mbulge = mstar_pri * mmb['bulge_frac']
    mbh = mmb['mamp'] * np.power(10.0, np.log10(mbulge / mmb['mref']) * plaw)
    dmstar_dmbh_pri = (1.0 / mmb['bulge_frac']) * (mbulge / (plaw * mbh))
    qterm = (1.0 + mstar_rat) / (1.0 + mrat[None, :, None])
    dens = dens * ((mtot[:, None, None] / mstar_tot) * ((dmstar_dmbh_pri * qterm) / dqbh_dqgal))
    return dens, zprime, gmt_time


# ---------------------------------------------------------------------------
# Evolution (C restatement)
# ---------------------------------------------------------------------------
def find_2pwl_norm(time, mt, mr, sepa_init, rchar, g_in, g_out, nsteps):
    mt = _c(mt)
    mr = _c(mr)
    out = np.zeros(mt.size)
    lib().or_find_2pwl_norm(time, _p(mt), _p(mr), mt.size, sepa_init, rchar, g_in, g_out, nsteps, _p(out))
    return out


def dbn_gw(fobs_orb, nden, mtot, mrat, redz_prime, cosmo):
    M, Q, Z = nden.shape
    F = len(fobs_orb)
    rf = -np.ones((M, Q, Z, F))
    dn = np.zeros((M, Q, Z, F))
    fo, nd, mt, mr, rp = [_c(a) for a in (fobs_orb, nden, mtot, mrat, redz_prime)]
    gz, gd = _c(cosmo._grid_z), _c(cosmo._grid_dcom)
    lib().or_dbn_gw(_p(fo), F, _p(nd), _p(mt), M, _p(mr), Q, Z, _p(rp), _p(gz), _p(gd), gz.size, _p(rf), _p(dn))
    return rf, dn


def dbn_2pwl(fobs_orb, sepa_init, num_steps, norm, rchar, g_in, g_out, nden, mtot, mrat, redz, gmt_time, cosmo):
    M, Q, Z = nden.shape
    F = len(fobs_orb)
    rf = -np.ones((M, Q, Z, F))
    dn = np.zeros((M, Q, Z, F))
    fo, nm, nd, mt, mr, rz, gt = [_c(a) for a in (fobs_orb, norm, nden, mtot, mrat, redz, gmt_time)]
    gz, gd, ga = _c(cosmo._grid_z), _c(cosmo._grid_dcom), _c(cosmo._grid_age)
    lib().or_dbn_2pwl(_p(fo), F, sepa_init, num_steps, _p(nm), rchar, g_in, g_out, _p(nd), _p(mt), M, _p(mr), Q,
                      _p(rz), Z, _p(gt), _p(gz), _p(gd), _p(ga), gz.size, _p(rf), _p(dn))
    return rf, dn


def integrate_3dx1d(edges, dnum):
    M, Q, Z, F = dnum.shape
    lm = _c(np.log10(edges[0]))
    mr = _c(edges[1])
    rz = _c(edges[2])
    dlnf = _c(np.diff(np.log(edges[3])))
    dn = _c(dnum)
    out = np.zeros((M - 1, Q - 1, Z - 1, F))
    lib().or_integrate_3dx1d(_p(lm), M, _p(mr), Q, _p(rz), Z, _p(dlnf), F, _p(dn), _p(out))
    return out


# ---------------------------------------------------------------------------
# Strain (gravwaves.py:694-725)
# ---------------------------------------------------------------------------
def midpoints(vals, axis=-1):
    mm = np.moveaxis(np.asarray(vals), axis, 0)
    mm = 0.5 * (mm[1:] + mm[:-1])
    return np.moveaxis(mm, 0, axis)


def chirp_mass_mtmr(mt, mr):
    """utils.chirp_mass_mtmr (utils.py:1978-1997)."""
    return mt * np.power(mr, 3.0 / 5.0) / np.power(1 + mr, 6.0 / 5.0)


def gw_strain_source(mchirp, dcom, freq_rest_orb):
    """utils.gw_strain_source (utils.py:2260-2285)."""
    return _GW_SRC_CONST * mchirp * np.power(2 * mchirp * freq_rest_orb, 2 / 3) / dcom


def h2fdf_redz(edges, redz, cosmo):
    fo = np.asarray(edges[-1])
    df = np.diff(fo)
    fc = midpoints(fo)
    for dd in range(3):
        redz = midpoints(redz, axis=dd)
    mt = midpoints(edges[0])
    mr = midpoints(edges[1])
    mc = chirp_mass_mtmr(mt[:, None], mr[None, :])
    mc = mc[:, :, None, None]
    dc = np.inf * np.ones_like(redz)
    sel = redz > 0.0
    dc[sel] = cosmo.comoving_distance(redz[sel])
    fr = fc[None, None, None, :] * (1.0 + redz)
    hs = gw_strain_source(mc, dc, fr)
    return (hs ** 2) * (fc / df)


def bin_params_redz(edges, redz, cosmo):
    """(7, B, F) [mt, mr, rz, redz_final, dcom_final, sepa, angs] (single_sources.py:104-139)."""
    for dd in range(3):
        redz = midpoints(redz, axis=dd)
    redz = np.array(redz)
    dcom = np.inf * np.ones_like(redz)
    sel = redz > 0.0
    redz[~sel] = -1.0
    redz[redz < 0] = -1.0
    dcom[sel] = cosmo.comoving_distance(redz[sel])
    fc = midpoints(np.asarray(edges[-1]))
    mt = midpoints(edges[0])
    mr = midpoints(edges[1])
    rz = midpoints(edges[2])
    frst = fc[None, None, None, :] * (1.0 + redz)
    with np.errstate(divide='ignore', invalid='ignore'):
        sepa = np.power(NWTG * mt[:, None, None, None] / np.square(2.0 * np.pi * frst), 1.0 / 3.0)
        angs = sepa / (dcom / (1.0 + redz))
    shp = redz.shape
    full = [np.broadcast_to(mt[:, None, None, None], shp), np.broadcast_to(mr[None, :, None, None], shp),
            np.broadcast_to(rz[None, None, :, None], shp), redz, dcom, sepa, angs]
    F = shp[-1]
    return np.stack([np.asarray(a).reshape(-1, F) for a in full])


# ---------------------------------------------------------------------------
# Realization walks (C restatement) with injected counts
# ---------------------------------------------------------------------------
def loudest_counts(counts, h2fdf, order, L, skip_zero_h=False, par=None):
    """counts (B,F,R), h2fdf (B,F), order (B,) -> hc2ss (F,R,L), ssidx (F,R,L), hc2bg (F,R), bgpar."""
    B, F, R = counts.shape
    cn, hh = _c(counts), _c(h2fdf)
    od = _c(order, np.int64)
    hc2ss = np.zeros((F, R, L))
    ssidx = -np.ones((F, R, L), dtype=np.int64)
    hc2bg = np.zeros((F, R))
    npar = 0 if par is None else par.shape[0]
    pp = None if par is None else _c(par)
    bgpar = np.zeros((npar, F, R)) if npar else None
    lib().or_loudest_counts(_p(cn), _p(hh), _p(od), B, F, R, L, int(skip_zero_h), _p(pp), npar, _p(hc2ss), _p(hc2bg),
                            _p(ssidx), _p(bgpar))
    return hc2ss, ssidx, hc2bg, bgpar


def ss_bg_counts(counts, h2fdf, par=None):
    """Per-frequency argmax split (cyutils.pyx:935-1178). par (npar, B)."""
    B, F, R = counts.shape
    cn, hh = _c(counts), _c(h2fdf)
    hc2ss = np.zeros((F, R))
    hc2bg = np.zeros((F, R))
    ssidx = np.zeros((F, R), dtype=np.int64)
    npar = 0 if par is None else par.shape[0]
    pp = None if par is None else _c(par)
    bgpar = np.zeros((npar, F, R)) if npar else None
    lib().or_ss_bg_counts(_p(cn), _p(hh), B, F, R, _p(pp), npar, _p(hc2ss), _p(hc2bg), _p(ssidx), _p(bgpar))
    return hc2ss, hc2bg, ssidx, bgpar


def poisson_gwb_counts(counts, h2fdf):
    B, F, R = counts.shape
    out = np.zeros((F, R))
    lib().or_poisson_gwb_counts(_p(_c(counts)), _p(_c(h2fdf)), B, F, R, _p(out))
    return out


def draw_counts(number, R, seed, thresh=1e10):
    B, F = number.shape
    out = np.zeros((B, F, R))
    lib().or_draw_counts(_p(_c(number)), B, F, R, thresh, seed, _p(out))
    return out


def loudest_rng(number, h2fdf, order, R, L, seed, thresh=1e10):
    B, F = number.shape
    hc2ss = np.zeros((F, R, L))
    hc2bg = np.zeros((F, R))
    lib().or_loudest_rng(_p(_c(number)), _p(_c(h2fdf)), _p(_c(order, np.int64)), B, F, R, L, thresh, seed,
                         _p(hc2ss), _p(hc2bg))
    return hc2ss, hc2bg


def order_desc(h2fdf0):
    """Stable descending order (the build's definition of the reference's argsort, single_sources.py:89)."""
    return np.argsort(-np.asarray(h2fdf0), kind='stable').astype(np.int64)


# ---------------------------------------------------------------------------
# Full pipeline from a configuration dict
# ---------------------------------------------------------------------------
def edges_from_cfg(cfg):
    M, Q, Z = cfg['shape']
    mtot = np.logspace(*np.log10(cfg['mtot']), M)
    mrat = np.logspace(*np.log10(cfg['mrat']), Q)
    redz = np.logspace(*np.log10(cfg['redz']), Z)
    return mtot, mrat, redz


def number_h2fdf(cfg, density=None):
    """Oracle stage A: returns dict with dens, zprime, gmt_time, norm, redz_final, dnum, number, h2fdf."""
    c = cfg['cosmo']
    cosmo = OracleCosmo(h=c['h'], Om0=c['Om0'], size=c.get('size', 200), zmax=c.get('zmax', 20.0))
    mtot, mrat, redz = edges_from_cfg(cfg)
    dens, zprime, gmt = static_density(mtot, mrat, redz, cfg, cosmo)
    if density is not None:
        dens = density          # an injected density
    elif cfg['mmbulge'].get('scatter_dex', 0.0) > 0.0:
        from .scatter import add_scatter_to_masses
        dens = add_scatter_to_masses(mtot, mrat, dens, cfg['mmbulge']['scatter_dex'])
    if zprime is not None:
        dens = np.array(dens)
        dens[zprime < 0.0] = 0.0
    fge = np.asarray(cfg['fobs_gw_edges'])
    fgc = midpoints(fge)
    foe, foc = fge / 2.0, fgc / 2.0
    hard = cfg['hard']
    out = dict(cosmo=cosmo, dens=dens, zprime=zprime, gmt_time=gmt, mtot=mtot, mrat=mrat, redz=redz)
    if hard['kind'] == 'gw':
        rp = zprime if zprime is not None else np.broadcast_to(redz[None, None, :], dens.shape)
        rf, dn = dbn_gw(foc, dens, mtot, mrat, rp, cosmo)
    else:
        mt2, mr2 = np.meshgrid(mtot, mrat, indexing='ij')
        nl = find_2pwl_norm(hard['time'], mt2.ravel(), mr2.ravel(), hard['sepa_init'], hard['rchar'],
                            hard['gamma_inner'], hard['gamma_outer'], hard['num_steps']).reshape(mt2.shape)
        norm = 10.0 ** nl
        out['norm_log10'] = nl
        gt = gmt if gmt is not None else np.zeros(dens.shape)
        rf, dn = dbn_2pwl(foc, hard['sepa_init'], hard['num_steps'], norm, hard['rchar'], hard['gamma_inner'],
                          hard['gamma_outer'], dens, mtot, mrat, redz, gt, cosmo)
    edges = [mtot, mrat, redz, foe]
    number = integrate_3dx1d(edges, dn)
    h2 = h2fdf_redz(edges, rf, cosmo)
    out.update(redz_final=rf, dnum=dn, number=number, h2fdf=h2, edges=edges)
    return out
