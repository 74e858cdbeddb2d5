"""Device-level pipeline of the SAM GWB hot path.

Every function takes and returns device tensors (float64 unless stated) and
issues its work on the current stream through the C-ABI of libholohip.so;
nothing here computes per-cell values on the host.
"""
import numpy as np
import torch

from . import _lib
from . import device as D
from .constants import NWTG, SPLC

#: 8 G^(5/3) pi^(2/3) / sqrt(10) / c^4 with astropy G, c (utils.py:40)
GW_SRC_CONST = 8 * np.power(NWTG, 5 / 3) * np.power(np.pi, 2 / 3) / np.sqrt(10) / np.power(SPLC, 4)
NORMAL_THRESHOLD = 1e10

#: when a dict, each stage records (start, end) HIP events on the current stream under its name
EVENTS = None


class _stage:
    def __init__(self, name):
        self.name = name

    def __enter__(self):
        if EVENTS is not None:
            self.e0 = torch.cuda.Event(enable_timing=True)
            self.e0.record()
        return self

    def __exit__(self, *exc):
        if EVENTS is not None:
            e1 = torch.cuda.Event(enable_timing=True)
            e1.record()
            EVENTS.setdefault(self.name, []).append((self.e0, e1))
        return False


def stage_times_ms():
    """Mean elapsed ms per recorded stage (synchronizes)."""
    torch.cuda.synchronize()
    return {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in (EVENTS or {}).items()}


def find_2pwl_norm(target_time, mtot, mrat, sepa_init, rchar, gamma_inner, gamma_outer, nsteps):
    """(log10 norm, norm) per element (sam_cyutils.pyx:289-353, hardening.py:1413-1416)."""
    mt = D.to_dev(mtot)
    mr = D.to_dev(mrat)
    out = D.empty(mt.numel())
    norm = D.empty(mt.numel())
    with _stage("norm_2pwl"):
        _lib.call("hd_find_2pwl_norm", float(target_time), D.ptr(mt), D.ptr(mr), mt.numel(), float(sepa_init),
                  float(rchar), float(gamma_inner), float(gamma_outer), int(nsteps), D.ptr(out), D.ptr(norm),
                  D.stream_ptr())
    return out, norm


def static_density(mtot_d, mrat_d, redz_d, par, cosmo):
    """(dens, zprime, gmt_time) before scatter and stalled-zeroing, all (M,Q,Z)."""
    M, Q, Z = mtot_d.numel(), mrat_d.numel(), redz_d.numel()
    tabs = cosmo.device_tables()
    dens = D.empty((M, Q, Z))
    zp = D.empty((M, Q, Z))
    gmt = D.empty((M, Q, Z))
    par = np.ascontiguousarray(par, dtype=np.float64)
    assert par.size == _lib.HD_DENS_NPAR
    with _stage("density"):
        _lib.call("hd_static_density", D.ptr(mtot_d), M, D.ptr(mrat_d), Q, D.ptr(redz_d), Z, par.ctypes.data,
                  D.ptr(tabs["dc_tab"]), D.ptr(tabs["dc_der"]), tabs["n_tab"], D.ptr(dens), D.ptr(zp), D.ptr(gmt),
                  D.stream_ptr())
    return dens, zp, gmt


def zero_stalled(dens, zprime):
    _lib.call("hd_zero_stalled", D.ptr(dens), D.ptr(zprime), dens.numel(), D.stream_ptr())


def dbn_gw(fobs_orb_d, nden, mtot_d, mrat_d, redz_prime, cosmo):
    M, Q, Z = nden.shape
    F = fobs_orb_d.numel()
    tabs = cosmo.device_tables()
    rf = D.empty((M, Q, Z, F))
    dn = D.empty((M, Q, Z, F))
    with _stage("dbn_gw"):
        _lib.call("hd_dbn_gw", D.ptr(fobs_orb_d), F, D.ptr(nden), D.ptr(mtot_d), M, D.ptr(mrat_d), Q, Z,
                  D.ptr(redz_prime), D.ptr(tabs["grid_z"]), D.ptr(tabs["grid_dcom"]), cosmo.size, D.ptr(rf), D.ptr(dn),
                  D.stream_ptr())
    return rf, dn


def dbn_2pwl(fobs_orb_d, fobs_sorted, sepa_init, num_steps, norm_d, rchar, gamma_inner, gamma_outer, nden, mtot_d,
             mrat_d, redz_d, gmt_time, cosmo):
    M, Q, Z = nden.shape
    F = fobs_orb_d.numel()
    tabs = cosmo.device_tables()
    rf = D.empty((M, Q, Z, F))
    dn = D.empty((M, Q, Z, F))
    with _stage("dbn_2pwl"):
        _lib.call("hd_dbn_2pwl", D.ptr(fobs_orb_d), F, int(bool(fobs_sorted)), float(sepa_init), int(num_steps),
                  D.ptr(norm_d), float(rchar), float(gamma_inner), float(gamma_outer), D.ptr(nden), D.ptr(mtot_d), M,
                  D.ptr(mrat_d), Q, D.ptr(redz_d), Z, D.ptr(gmt_time), D.ptr(tabs["grid_z"]), D.ptr(tabs["grid_dcom"]),
                  D.ptr(tabs["grid_age"]), cosmo.size, D.ptr(rf), D.ptr(dn), D.stream_ptr())
    return rf, dn


def integrate_3dx1d(log10_mtot_d, mrat_d, redz_d, dlnf_d, dnum):
    M, Q, Z, F = dnum.shape
    out = D.empty((M - 1, Q - 1, Z - 1, F))
    with _stage("integrate"):
        _lib.call("hd_integrate_3dx1d", D.ptr(log10_mtot_d), M, D.ptr(mrat_d), Q, D.ptr(redz_d), Z, D.ptr(dlnf_d), F,
                  D.ptr(dnum), D.ptr(out), D.stream_ptr())
    return out


def h2fdf_redz(mt_cent_d, mr_cent_d, fc_d, fcdf_d, redz_final, cosmo):
    M, Q, Z, F = redz_final.shape
    tabs = cosmo.device_tables()
    out = D.empty((M - 1, Q - 1, Z - 1, F))
    with _stage("h2fdf"):
        _lib.call("hd_h2fdf_redz", D.ptr(mt_cent_d), M, D.ptr(mr_cent_d), Q, Z, D.ptr(fc_d), D.ptr(fcdf_d), F,
                  D.ptr(redz_final), float(GW_SRC_CONST), cosmo.Om0, cosmo.h, D.ptr(tabs["dc_tab"]),
                  D.ptr(tabs["dc_der"]), tabs["n_tab"], D.ptr(out), D.stream_ptr())
    return out


def order_desc(key2d, col=0):
    """Stable descending order of rows of key2d (B, stride) by column `col` (int64 device tensor)."""
    B, stride = key2d.shape
    order = D.empty(B, dtype=torch.int64)
    ws = D.workspace(_lib.size("hd_order_workspace_size", B))
    with _stage("order"):
        _lib.call("hd_order_desc", D.ptr(key2d), B, stride, int(col), D.ptr(order), D.ptr(ws), ws.numel(),
                  D.stream_ptr())
    return order


def realize(number2d, h2fdf2d, nreals, nloud, mode, order=None, counts=None, flags=0, par=None, seed=0,
            normal_threshold=NORMAL_THRESHOLD, record_counts=None):
    """Realizations over (B,F) bins.  Returns (hc2ss (F,R,L), ssidx (F,R,L), hc2bg (F,R), bgpar (npar,F,R)|None).

    ``record_counts``: a (B, F, R) float64 device tensor that receives every count the draw kernels
    draw (the recording test hook; same outputs as without it for the same seed)."""
    B, F = h2fdf2d.shape
    R = int(nreals)
    L = 0 if mode == _lib.HD_SEL_NONE else int(nloud)
    npar = 0 if par is None else int(par.shape[0])
    hc2bg = D.empty((F, R))
    hc2ss = D.empty((F, R, max(L, 1)))
    ssidx = D.empty((F, R, max(L, 1)), dtype=torch.int64)
    bgpar = D.empty((npar, F, R)) if npar else None
    with _stage("realize"):
        if counts is not None:
            ws = D.workspace(_lib.size("hd_realize_counts_workspace_size", B, F, R, L, npar))
            _lib.call("hd_realize_counts", D.ptr(counts), D.ptr(h2fdf2d), D.ptr(order), B, F, R, L, mode, flags,
                      D.ptr(par), npar, D.ptr(hc2ss), D.ptr(ssidx), D.ptr(hc2bg), D.ptr(bgpar), D.ptr(ws), ws.numel(),
                      D.stream_ptr())
        elif record_counts is not None:
            if tuple(record_counts.shape) != (B, F, R) or record_counts.dtype != torch.float64:
                raise ValueError(f"record_counts must be float64 of shape {(B, F, R)}")
            ws = D.workspace(_lib.size("hd_realize_workspace_size", B, F, R, L, npar))
            _lib.call("hd_realize_record_counts", D.ptr(number2d), D.ptr(h2fdf2d), D.ptr(order), B, F, R, L, mode,
                      flags, float(normal_threshold), int(seed) & 0xFFFFFFFFFFFFFFFF, D.ptr(par), npar, D.ptr(hc2ss),
                      D.ptr(ssidx), D.ptr(hc2bg), D.ptr(bgpar), D.ptr(record_counts), D.ptr(ws), ws.numel(),
                      D.stream_ptr())
        else:
            ws = D.workspace(_lib.size("hd_realize_workspace_size", B, F, R, L, npar))
            _lib.call("hd_realize", D.ptr(number2d), D.ptr(h2fdf2d), D.ptr(order), B, F, R, L, mode, flags,
                      float(normal_threshold), int(seed) & 0xFFFFFFFFFFFFFFFF, D.ptr(par), npar, D.ptr(hc2ss),
                      D.ptr(ssidx), D.ptr(hc2bg), D.ptr(bgpar), D.ptr(ws), ws.numel(), D.stream_ptr())
    if L == 0:
        hc2ss = hc2ss[..., :0]
        ssidx = ssidx[..., :0]
    return hc2ss, ssidx, hc2bg, bgpar


# ---- eccentric SAM (cyutils.pyx:361-851) ----
ECCEN_WS_COLUMNS_BYTES = 8 << 30    # device bytes budget of one harmonic chunk of the discrete variant


def _eccen_args(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d):
    M, Q, Z = ndens.shape
    for t, n in ((mtot_log10_d, M), (mrat_d, Q), (redz_d, Z), (dcom_d, Z)):
        if t.numel() != n:
            raise ValueError("grid arrays must match the (M, Q, Z) density")
    if sepa_d.numel() != eccen_d.numel() or sepa_d.numel() < 2:
        raise ValueError("sepa_evo and eccen_evo must have the same length >= 2")
    return M, Q, Z, gwfobs_d.numel(), sepa_d.numel()


def gwb_eccen(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d, nharms):
    """(F, nharms) characteristic strain squared per frequency and harmonic (device)."""
    M, Q, Z, F, E = _eccen_args(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d)
    ws = D.workspace(_lib.size("hd_eccen_workspace_size", M, Z, E))
    out = D.empty((F, int(nharms)))
    with _stage("eccen"):
        _lib.call("hd_gwb_eccen", D.ptr(ndens), D.ptr(mtot_log10_d), M, D.ptr(mrat_d), Q, D.ptr(redz_d),
                  D.ptr(dcom_d), Z, D.ptr(gwfobs_d), F, D.ptr(sepa_d), D.ptr(eccen_d), E, int(nharms), D.ptr(out),
                  D.ptr(ws), ws.numel(), D.stream_ptr())
    return out


def eccen_number(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d, h0, hc):
    """Poisson means and strain factors (M*Q*Z, F*hc) of harmonics h0+1 .. h0+hc (device)."""
    M, Q, Z, F, E = _eccen_args(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d)
    ws = D.workspace(_lib.size("hd_eccen_workspace_size", M, Z, E))
    lam = D.empty((M * Q * Z, F * int(hc)))
    hv = D.empty((M * Q * Z, F * int(hc)))
    with _stage("eccen_number"):
        _lib.call("hd_eccen_number", D.ptr(ndens), D.ptr(mtot_log10_d), M, D.ptr(mrat_d), Q, D.ptr(redz_d),
                  D.ptr(dcom_d), Z, D.ptr(gwfobs_d), F, D.ptr(sepa_d), D.ptr(eccen_d), E, int(h0), int(hc),
                  D.ptr(lam), D.ptr(hv), D.ptr(ws), ws.numel(), D.stream_ptr())
    return lam, hv


def eccen_harm_chunk(B, F, nharms, nreals):
    """Harmonics per chunk of the discrete variant: Poisson means, strain factors and the
    realization workspace of one chunk within ECCEN_WS_COLUMNS_BYTES."""
    per_col = 16 * B + _lib.size("hd_realize_workspace_size", B, F, int(nreals), 0, 0) / max(F, 1)
    return int(max(1, min(nharms, ECCEN_WS_COLUMNS_BYTES // (per_col * F))))


def gwb_eccen_discrete(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d, nharms, nreals, seed=0,
                       harm_chunk=None, normal_threshold=float("inf")):
    """(F, nharms, R) characteristic strain squared with a Poisson count per (node, frequency,
    harmonic, realization) (device).  Every count is Poisson, as the reference's random_poisson
    call (cyutils.pyx:834) -- no Normal approximation unless ``normal_threshold`` is given.
    Harmonics run in chunks; chunk c draws with the Philox seed seed + (c+1) * 0x9E3779B97F4A7C15,
    so a result depends on (seed, harm_chunk)."""
    M, Q, Z, F, E = _eccen_args(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d)
    H, R = int(nharms), int(nreals)
    hc = int(harm_chunk) if harm_chunk else eccen_harm_chunk(M * Q * Z, F, H, R)
    out = D.empty((F, H, R))
    for c, h0 in enumerate(range(0, H, hc)):
        n = min(hc, H - h0)
        lam, hv = eccen_number(ndens, mtot_log10_d, mrat_d, redz_d, dcom_d, gwfobs_d, sepa_d, eccen_d, h0, n)
        s = (int(seed) + (c + 1) * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        _, _, hc2, _ = realize(lam, hv, R, 0, _lib.HD_SEL_NONE, seed=s, normal_threshold=normal_threshold)
        out[:, h0:h0 + n, :] = hc2.view(F, n, R)
        del lam, hv, hc2
    return out


def mass_stellar(mtot_d, M, mrat_d, Q, Z, par, mpri, mrat_out, mtot_out):
    """Stellar masses per (M, Q, Z) node into the three device arrays (sam.py:250-278)."""
    par = np.ascontiguousarray(par, dtype=np.float64)
    _lib.call("hd_mass_stellar", D.ptr(mtot_d), M, D.ptr(mrat_d), Q, Z, par.ctypes.data, D.ptr(mpri), D.ptr(mrat_out),
              D.ptr(mtot_out), D.stream_ptr())
