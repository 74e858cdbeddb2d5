"""Semi-analytic model (SAM) of MBH binary populations -- the GWB hot path.

Mirrors ``holodeck.sams.sam.Semi_Analytic_Model`` (holodeck/sams/sam.py:67-947):
same constructor, grid edges, ``static_binary_density``,
``dynamic_binary_number_at_fobs`` and ``gwb``.  Every per-cell quantity is
computed on the device through libholohip.so; the host handles only the 1-D
grid edges, the frequency bins and the final (F, R[, L]) square roots.

Extensions (keyword-only, absent from the reference):
* ``gwb(..., seed=...)`` seeds the counter-based generator (the reference's
  realizations use an unseeded PCG64 per call, cyutils.pyx:1313-1317);
* ``gwb(..., counts=...)`` injects the per-(bin, frequency, realization) counts
  instead of drawing them (exact, deterministic comparison path);
* ``gwb(..., realize=False)`` returns expectation values: hc_bg**2 =
  sum(number * h2fdf) as gravwaves.py:509-514, hc_ss from floor(number)
  counts; the reference's ``gwb`` raises for non-integer ``realize``.
* ``gwb(..., loud_mode='argmax')`` selects the per-frequency loudest
  instances (ss_bg_hc semantics) instead of the frequency-0 order walk.
"""
import numpy as np

from .. import _lib
from .. import device as D
from .. import utils
from .. import host_relations
from ..constants import MSOL, SPLC, MPC, NWTG
from ..cosmology import cosmo as _default_cosmo
from .components import (
    _Galaxy_Pair_Fraction, _Galaxy_Stellar_Mass_Function, _Galaxy_Merger_Time, _Galaxy_Merger_Rate,
    GSMF_Schechter, GPF_Power_Law, GMT_Power_Law, GMR_Illustris, density_params,
)

GSMF_USES_MTOT = False
GPF_USES_MTOT = False
GMT_USES_MTOT = False


class Semi_Analytic_Model:
    """SAM over a rectilinear (M, q, z) grid of edges (sam.py:67-234)."""

    def __init__(self, mtot=(1.0e4 * MSOL, 1.0e12 * MSOL, 91), mrat=(1e-3, 1.0, 81), redz=(1e-3, 10.0, 101),
                 shape=None, log=None, gsmf=GSMF_Schechter, gpf=None, gmt=None, gmr=None,
                 mmbulge=host_relations.MMBulge_KH2013, cosmo=None, **kwargs):
        for key in kwargs:
            raise ValueError(f"Unexpected kwarg {key=}: {kwargs[key]=}!")
        self._cosmo = _default_cosmo if cosmo is None else cosmo
        gsmf = utils.get_subclass_instance(gsmf, None, _Galaxy_Stellar_Mass_Function)
        mmbulge = utils.get_subclass_instance(mmbulge, None, host_relations._MMBulge_Relation)
        if gpf is None:
            gmr = utils.get_subclass_instance(gmr, GMR_Illustris, _Galaxy_Merger_Rate)
            gmt = utils.get_subclass_instance(gmt, None, _Galaxy_Merger_Time, allow_none=True)
        else:
            if gmr is not None:
                raise ValueError("Can only use one of `gpf` and `gmr`!")
            gmt = utils.get_subclass_instance(gmt, GMT_Power_Law, _Galaxy_Merger_Time)
            gpf = utils.get_subclass_instance(gpf, GPF_Power_Law, _Galaxy_Pair_Fraction)
        self._gsmf = gsmf
        self._gmr = gmr
        self._gpf = gpf
        self._gmt = gmt
        self._mmbulge = mmbulge

        if shape is not None and np.isscalar(shape):
            shape = [shape for ii in range(3)]
        params = [mtot, mrat, redz]
        names = ['mtot', 'mrat', 'redz']
        for ii, (par, name) in enumerate(zip(params, names)):
            if not isinstance(par, tuple) and (len(par) == 3):
                raise ValueError(f"{name} must be a (3,) tuple specifying a log-spacing!")
            par = [pp for pp in par]
            if shape is not None and shape[ii] is not None:
                par[2] = shape[ii]
            params[ii] = np.logspace(*np.log10(par[:2]), par[2])
        self.mtot, self.mrat, self.redz = params

        self._density = None
        self._shape = None
        self._dev = None          # device copies of grids and density products
        self._scatter_iters = None

    # ---- grid ----
    @property
    def edges(self):
        return [self.mtot, self.mrat, self.redz]

    @property
    def shape(self):
        if self._shape is None:
            self._shape = tuple([len(ee) for ee in self.edges])
        return self._shape

    @property
    def cosmo(self):
        return self._cosmo

    def _grids(self):
        if self._dev is None:
            self._dev = dict(mtot=D.to_dev(self.mtot), mrat=D.to_dev(self.mrat), redz=D.to_dev(self.redz),
                             log10_mtot=D.to_dev(np.log10(self.mtot)),
                             mt_cent=D.to_dev(utils.midpoints(self.mtot)),
                             mr_cent=D.to_dev(utils.midpoints(self.mrat)))
        return self._dev

    def mass_stellar(self):
        """(mstar_pri, mstar_rat, mstar_tot, redz), each (M, Q, Z), from the M-Mbulge relation
        without scatter (sam.py:250-278); computed on the device, returned as numpy."""
        from .. import engine
        g = self._grids()
        par = density_params(self._gsmf, self._gpf, self._gmt, self._gmr, self._mmbulge, self._cosmo)
        M, Q, Z = self.shape
        out = [D.empty((M, Q, Z)) for _ in range(3)]
        engine.mass_stellar(g['mtot'], M, g['mrat'], Q, Z, par, *out)
        mp, mr, mt = (D.to_host(t) for t in out)
        redz = np.broadcast_to(self.redz[np.newaxis, np.newaxis, :], (M, Q, Z))
        return mp, mr, mt, redz

    # ---- density ----
    def _density_device(self):
        """Device (dens, zprime|None, gmt_time|None): static_binary_density (sam.py:280-398)."""
        g = self._grids()
        if 'dens' in g:
            return g['dens'], g['zprime'], g['gmt_time']
        from .. import engine
        par = density_params(self._gsmf, self._gpf, self._gmt, self._gmr, self._mmbulge, self._cosmo)
        dens, zp, gmt = engine.static_density(g['mtot'], g['mrat'], g['redz'], par, self._cosmo)
        scatter = self._mmbulge._scatter_dex
        if scatter > 0.0:
            from .. import scatter as _scatter
            self._scatter_iters = _scatter.add_scatter_to_masses_device(self.mtot, self.mrat, dens, scatter)
        if self._gmt is not None:
            engine.zero_stalled(dens, zp)
            g['zprime'] = zp
            g['gmt_time'] = gmt
        else:
            g['zprime'] = None
            g['gmt_time'] = None
        g['dens'] = dens
        return g['dens'], g['zprime'], g['gmt_time']

    @property
    def static_binary_density(self):
        """d^3 n / [dlog10 M dq dz] at the grid edges [Mpc^-3], (M, Q, Z) (sam.py:280-398)."""
        if self._density is None:
            dens, _, _ = self._density_device()
            self._density = D.to_host(dens)
            self._check_scatter()
        return self._density

    @property
    def _redz_prime(self):
        zp = self._density_device()[1]
        return None if zp is None else D.to_host(zp)

    @property
    def _gmt_time(self):
        gt = self._density_device()[2]
        return None if gt is None else D.to_host(gt)

    # ---- evolution ----
    def _dbn_device(self, hard, fobs_orb):
        """Device (redz_final, diff_num) (M,Q,Z,F): sam_cyutils.dynamic_binary_number_at_fobs (sam_cyutils.pyx:421-504)."""
        from .. import engine
        from .. import hardening
        fobs_orb = np.atleast_1d(np.asarray(fobs_orb, dtype=float))
        g = self._grids()
        nden, zp, gmt = self._density_device()
        fo = D.to_dev(fobs_orb)
        if isinstance(hard, hardening.Fixed_Time_2PL_SAM):
            if gmt is None:
                gmt = D.empty(self.shape)
                gmt.zero_()
            srt = bool(np.all(np.diff(fobs_orb) >= 0))
            return engine.dbn_2pwl(fo, srt, hard._sepa_init, hard._num_steps, hard._norm_dev, hard._rchar,
                                   hard._gamma_inner, hard._gamma_outer, nden, g['mtot'], g['mrat'], g['redz'], gmt,
                                   self._cosmo)
        if isinstance(hard, hardening.Hard_GW) or (isinstance(hard, type) and issubclass(hard, hardening.Hard_GW)):
            if zp is None:
                zp = g['redz'].reshape(1, 1, -1).expand(*self.shape).contiguous()
            return engine.dbn_gw(fo, nden, g['mtot'], g['mrat'], zp, self._cosmo)
        raise ValueError(f"Unexpected `hard` value {hard}!")

    def dynamic_binary_number_at_fobs(self, hard, fobs_orb, use_cython=True, **kwargs):
        """(grid, dnum, redz_final) at observer-frame orbital frequencies (sam.py:400-466)."""
        rf, dn = self._dbn_device(hard, fobs_orb)
        grid = [self.mtot, self.mrat, self.redz, np.asarray(fobs_orb)]
        return grid, D.to_host(dn), D.to_host(rf)

    # ---- GWB ----
    def _number_h2fdf_device(self, fobs_gw_edges, hard):
        """number (B,F), h2fdf (B,F), redz_final (M,Q,Z,F) on the device for GW-frequency bin edges."""
        from .. import engine
        fobs_gw_edges = np.asarray(fobs_gw_edges, dtype=float)
        fobs_gw_cents = utils.midpoints(fobs_gw_edges)
        fobs_orb_edges = fobs_gw_edges / 2.0
        fobs_orb_cents = fobs_gw_cents / 2.0
        rf, dn = self._dbn_device(hard, fobs_orb_cents)
        g = self._grids()
        dlnf = D.to_dev(np.diff(np.log(fobs_orb_edges)))
        number = engine.integrate_3dx1d(g['log10_mtot'], g['mrat'], g['redz'], dlnf, dn)
        fc = utils.midpoints(fobs_orb_edges)
        df = np.diff(fobs_orb_edges)
        h2fdf = engine.h2fdf_redz(g['mt_cent'], g['mr_cent'], D.to_dev(fc), D.to_dev(fc / df), rf, self._cosmo)
        F = fc.size
        return number.view(-1, F), h2fdf.view(-1, F), rf

    def gwb(self, fobs_gw_edges, hard=None, realize=100, loudest=1, params=False, *, seed=None, counts=None,
            loud_mode='sorted'):
        """GWB and loudest single sources at the given GW-frequency bin edges (sam.py:872-947).

        Returns ``(hc_ss (F,R,L), hc_bg (F,R))`` or, with ``params=True``,
        ``(hc_ss, hc_bg, sspar (4,F,R,L), bgpar (7,F,R))``.
        """
        from .. import hardening, single_sources
        if hard is None:
            hard = hardening.Hard_GW()
        if not isinstance(hard, (hardening.Fixed_Time_2PL_SAM, hardening.Hard_GW)):
            raise ValueError("`sam_cyutils` methods only work with `Fixed_Time_2PL_SAM` or `Hard_GW` hardening "
                             "models!")
        number, h2fdf, rf = self._number_h2fdf_device(fobs_gw_edges, hard)
        edges = [self.mtot, self.mrat, self.redz, np.asarray(fobs_gw_edges, dtype=float) / 2.0]
        out = single_sources._ss_gws_redz_device(edges, rf, number, h2fdf, realize=realize, loudest=loudest,
                                                 params=params, seed=seed, counts=counts, loud_mode=loud_mode,
                                                 cosmo=self._cosmo)
        self._check_scatter()
        return out

    def gwb_device(self, fobs_gw_edges, hard, realize, loudest, seed=0):
        """Device-resident gwb: (hc2ss (F,R,L), ssidx (F,R,L), hc2bg (F,R)) tensors, no host copies.

        Same pipeline as ``gwb(..., params=False)`` (frequency-0 order walk), for
        callers that keep results in HBM (batch drivers, benchmarks)."""
        from .. import engine
        number, h2fdf, rf = self._number_h2fdf_device(fobs_gw_edges, hard)
        order = engine.order_desc(h2fdf, col=0)
        hc2ss, ssidx, hc2bg, _ = engine.realize(number, h2fdf, int(realize), int(loudest), _lib.HD_SEL_SORTED,
                                                order=order, seed=seed)
        return hc2ss, ssidx, hc2bg

    def _check_scatter(self):
        """Raise as the reference does (sam.py:1377-1383) if the nearest-node fill left bad values."""
        it = self._scatter_iters
        if it is not None:
            bad = int(it[-1].item())
            if bad & (1 << 21):
                raise ValueError("M-Mbulge scatter: a level of the sweep schedule has more vertices than the "
                                 "plan's max_level_width")
            if bad >= (1 << 20):
                raise ValueError("M-Mbulge scatter: the triangulation has more overflow adjacency entries than "
                                 "the plan's n_ovf")
            if bad:
                raise ValueError(f"After 0th order interpolation, {bad} bad values remain!")

    def gwb_ideal(self, fobs_gw, sum=True, redz_prime=None):
        """Idealised, continuous GWB h_c at GW frequencies (sam.py:837-870; Phinney 2001 Eq. 5):
        purely GW-driven, no finite-number effects, no coalescence cut.  With a GMT (or
        ``redz_prime=True``) the post-merger redshifts z' are used.  Returns (F,) or, with
        ``sum=False``, (M-1, Q-1, Z-1, F)."""
        from .. import gravwaves
        if redz_prime is None:
            redz_prime = (self._gmt is not None)
        elif redz_prime and (self._gmt is None):
            raise AttributeError("No `GMT` instance stored, cannot use `redz_prime` values!")
        dens, zp, _ = self._density_device()
        self._check_scatter()
        g = self._grids()
        rz = zp if redz_prime else g['redz']
        return gravwaves._gwb_ideal_device(fobs_gw, dens, g['log10_mtot'], g['mtot'], g['mrat'], rz, sum=sum)


def evolve_eccen_uniform_single(sam, eccen_init, sepa_init, nsteps):
    """Eccentricity along log-spaced separations from ``sepa_init`` [cm] down to three
    Schwarzschild radii of the lightest binary, for GW-driven evolution (sam.py:1235-1288).

    Classic fourth-order Runge-Kutta in separation of de/da = 1 / (da/de) (Peters 1964,
    Eq. 5.7), eccentricity floored at zero after each step.  Host code: E sequential steps.
    Returns (sepa (E,), eccen (E,)).
    """
    from ..constants import SCHW
    if not (0.0 <= eccen_init < 1.0):
        raise ValueError(f"eccen_init must be in [0, 1), got {eccen_init}")

    def de_da(a, e):
        e2 = e * e
        num = 1.0 + (73.0 / 24.0) * e2 + (37.0 / 96.0) * e2 * e2
        den = (1.0 - e2) * (1.0 + (121.0 / 304.0) * e2)
        with np.errstate(divide='ignore'):
            return 1.0 / ((12.0 / 19.0) * (a / e) * (num / den)) if e != 0.0 else 0.0

    a_min = (SCHW * np.asarray(sam.mtot) * 3).min()
    sepa = np.logspace(np.log10(sepa_init), np.log10(a_min), nsteps)
    eccen = np.zeros(nsteps)
    eccen[0] = eccen_init
    for s in range(1, nsteps):
        a0, e0 = sepa[s - 1], eccen[s - 1]
        h = sepa[s] - a0
        k1 = h * de_da(a0, e0)
        k2 = h * de_da(a0 + h / 2.0, e0 + k1 / 2.0)
        k3 = h * de_da(a0 + h / 2.0, e0 + k2 / 2.0)
        k4 = h * de_da(a0 + h, e0 + k3)
        eccen[s] = max(e0 + (k1 + 2 * k2 + 2 * k3 + k4) / 6.0, 0.0)
    return sepa, eccen
