"""Single-source / background split of the GWB (holodeck/single_sources.py:40-173, 387-516)."""
import numpy as np
import torch

from . import _lib
from . import device as D
from . import engine
from . import utils

par_names = np.array(['mtot', 'mrat', 'redz_init', 'redz_final', 'dcom_final', 'sepa_final', 'angs_final'])


def _unravel(idx, shape3):
    """flat bin index (any shape) -> (3, ...) (m, q, z), -1 preserved."""
    idx = np.asarray(idx)
    out = np.full((3,) + idx.shape, -1, dtype=np.int64)
    ok = idx >= 0
    mm, qq, zz = np.unravel_index(idx[ok], shape3)
    out[0][ok], out[1][ok], out[2][ok] = mm, qq, zz
    return out


def _ss_gws_redz_device(edges, redz_final, number, h2fdf, realize, loudest=1, params=False, seed=None, counts=None,
                        loud_mode='sorted', cosmo=None):
    """Core of ss_gws_redz on device arrays: redz_final (M,Q,Z,F), number/h2fdf (B,F)."""
    B, F = h2fdf.shape
    M, Q, Z = [len(ee) for ee in edges[:3]]
    shape3 = (M - 1, Q - 1, Z - 1)
    bad = ((redz_final < 0) & (redz_final != -1)).any()
    if bool(bad):
        raise ValueError("redz < 0 and !=-1 found in redz, in ss_gws_redz()")
    mode = _lib.HD_SEL_SORTED if loud_mode == 'sorted' else _lib.HD_SEL_ARGMAX
    order = engine.order_desc(h2fdf, col=0) if mode == _lib.HD_SEL_SORTED else None
    flags = _lib.HD_FLAG_SKIP_ZERO_H if params else 0

    par = None
    if params:
        par = bin_params_redz(edges, redz_final, cosmo)      # (7, B, F)

    if realize is False:
        cnt = torch.floor(number).reshape(B, F, 1).contiguous()
        hc2ss, ssidx, _, bgpar = engine.realize(number, h2fdf, 1, loudest, mode, order=order, counts=cnt,
                                                flags=flags, par=par)
        _, _, hc2bg, _ = engine.realize(number, h2fdf, 1, 0, _lib.HD_SEL_NONE,
                                        counts=number.reshape(B, F, 1).contiguous())
        R = 1
    elif counts is not None:
        cnt = D.to_dev(counts).reshape(B, F, -1).contiguous()
        R = cnt.shape[-1]
        hc2ss, ssidx, hc2bg, bgpar = engine.realize(number, h2fdf, R, loudest, mode, order=order, counts=cnt,
                                                    flags=flags, par=par)
    elif utils.isinteger(realize):
        R = int(realize)
        if seed is None:
            seed = int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
        hc2ss, ssidx, hc2bg, bgpar = engine.realize(number, h2fdf, R, loudest, mode, order=order, flags=flags,
                                                    par=par, seed=seed)
    else:
        raise Exception("`realize` ({}) must be an integer!".format(realize))

    hc_ss = np.sqrt(D.to_host(hc2ss))
    hc_bg = np.sqrt(D.to_host(hc2bg))
    if not params:
        return hc_ss, hc_bg
    # sspar (4, F, R, L): [mt, mr, rz] bin centres and redz_final of each loud source
    idx = D.to_host(ssidx)
    mqz = _unravel(idx, shape3)
    mt = utils.midpoints(edges[0])
    mr = utils.midpoints(edges[1])
    rz = utils.midpoints(edges[2])
    sspar = np.zeros((4,) + idx.shape)
    ok = idx >= 0
    sspar[0][ok] = mt[mqz[0][ok]]
    sspar[1][ok] = mr[mqz[1][ok]]
    sspar[2][ok] = rz[mqz[2][ok]]
    zf = par[3]                                             # (B, F) device
    fidx = np.broadcast_to(np.arange(F)[:, None, None], idx.shape)
    gath = zf[torch.as_tensor(np.where(ok, idx, 0), device=zf.device),
              torch.as_tensor(fidx.copy(), device=zf.device)]
    sspar[3][ok] = D.to_host(gath)[ok]
    bg = D.to_host(bgpar)
    return hc_ss, hc_bg, sspar, bg


def bin_params_redz(edges, redz_final, cosmo):
    """(7, B, F) device array of [mt, mr, rz, redz_final, dcom_final, sepa, angs] per bin and frequency,
    as single_sources.py:104-139 builds them."""
    from .cosmology import cosmo as _c
    cosmo = _c if cosmo is None else cosmo
    M, Q, Z, F = redz_final.shape
    tabs = cosmo.device_tables()
    mt = D.to_dev(utils.midpoints(edges[0]))
    mr = D.to_dev(utils.midpoints(edges[1]))
    rz = D.to_dev(utils.midpoints(edges[2]))
    fc = D.to_dev(utils.midpoints(edges[3]))
    out = D.empty((7, (M - 1) * (Q - 1) * (Z - 1), F))
    _lib.call("hd_bin_params_redz", D.ptr(mt), M, D.ptr(mr), Q, D.ptr(rz), Z, D.ptr(fc), F, D.ptr(redz_final),
              cosmo.Om0, cosmo.h, D.ptr(tabs["dc_tab"]), D.ptr(tabs["dc_der"]), tabs["n_tab"], D.ptr(out),
              D.stream_ptr())
    return out


def ss_gws_redz(edges, redz, number, realize, loudest=1, params=False, *, seed=None, counts=None,
                loud_mode='sorted', cosmo=None):
    """Loudest single sources and background from a number grid (single_sources.py:40-173).

    edges: [mtot (M), mrat (Q), redz (Z), fobs_orb_edges (F+1)]; redz (M,Q,Z,F); number (M-1,Q-1,Z-1,F).
    """
    from .cosmology import cosmo as _c
    cosmo = _c if cosmo is None else cosmo
    rf = D.to_dev(redz)
    M, Q, Z, F = rf.shape
    num = D.to_dev(number).reshape(-1, F).contiguous()
    fo = np.asarray(edges[3], dtype=float)
    fc = utils.midpoints(fo)
    df = np.diff(fo)
    h2 = engine.h2fdf_redz(D.to_dev(utils.midpoints(edges[0])), D.to_dev(utils.midpoints(edges[1])), D.to_dev(fc),
                           D.to_dev(fc / df), rf, cosmo).reshape(-1, F)
    return _ss_gws_redz_device(edges, rf, num, h2, realize, loudest=loudest, params=params, seed=seed,
                               counts=counts, loud_mode=loud_mode, cosmo=cosmo)
