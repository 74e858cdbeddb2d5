"""GW strain of a binned population (holodeck/gravwaves.py:470-542, 694-725), device compute."""
import numpy as np

from . import _lib
from . import device as D
from . import engine
from . import utils


def char_strain_sq_from_bin_edges_redz(edges, redz, cosmo=None):
    """h_s^2 f/df per bin: edges [mtot, mrat, redz, fobs_orb_edges], redz (M,Q,Z,F) -> (M-1,Q-1,Z-1,F)."""
    from .cosmology import cosmo as _c
    cosmo = _c if cosmo is None else cosmo
    assert len(edges) == 4
    fo = np.asarray(edges[-1], dtype=float)
    fc = utils.midpoints(fo)
    df = np.diff(fo)
    rf = D.to_dev(redz)
    out = engine.h2fdf_redz(D.to_dev(utils.midpoints(edges[0])), D.to_dev(utils.midpoints(edges[1])),
                            D.to_dev(fc), D.to_dev(fc / df), rf, cosmo)
    return D.to_host(out)


def _gws_from_number_grid_integrated_redz(edges, redz, number, realize, sum=True, *, seed=None, cosmo=None):
    """h_c of the GWB without a single-source split (gravwaves.py:470-542); realize int -> (F, R)."""
    from .cosmology import cosmo as _c
    cosmo = _c if cosmo is None else cosmo
    hc2 = char_strain_sq_from_bin_edges_redz(edges, redz, cosmo=cosmo)
    if realize in [None, False]:
        F = hc2.shape[-1]
        n = D.to_dev(number).reshape(-1, F).contiguous()
        h = D.to_dev(hc2).reshape(-1, F).contiguous()
        if not sum:
            return np.sqrt(hc2 * number)
        _, _, hb, _ = engine.realize(n, h, 1, 0, _lib.HD_SEL_NONE, counts=n.reshape(-1, F, 1).contiguous())
        return np.sqrt(D.to_host(hb)[:, 0])
    if utils.isinteger(realize):
        if not sum:
            raise NotImplementedError("per-bin realizations (sum=False) are not supported on the device path")
        from .cyutils import sam_poisson_gwb
        return np.sqrt(sam_poisson_gwb(number, hc2, realize, seed=seed))
    raise ValueError("`realize` ({}) must be one of {{True, False, integer}}!".format(realize))


def _eccen_inputs(gwfobs, sam, sepa_evo, eccen_evo):
    """Device inputs of the eccentric kernels for a SAM (gravwaves.py:936-941, 967-972)."""
    from .constants import MPC
    ndens, _, _ = sam._density_device()
    sam._check_scatter()
    dcom = sam.cosmo.comoving_distance(sam.redz) / MPC
    return (ndens, D.to_dev(np.log10(sam.mtot)), D.to_dev(sam.mrat), D.to_dev(sam.redz), D.to_dev(dcom),
            D.to_dev(np.atleast_1d(np.asarray(gwfobs, dtype=float))), D.to_dev(np.asarray(sepa_evo, dtype=float)),
            D.to_dev(np.asarray(eccen_evo, dtype=float)))


def sam_calc_gwb_single_eccen(gwfobs, sam, sepa_evo, eccen_evo, nharms=100):
    """(F, nharms) GWB strain squared per harmonic of an eccentric SAM (gravwaves.py:933-941);
    ``np.sqrt(gwb.sum(axis=1))`` is h_c(f)."""
    return D.to_host(engine.gwb_eccen(*_eccen_inputs(gwfobs, sam, sepa_evo, eccen_evo), int(nharms)))


def sam_calc_gwb_single_eccen_discrete(gwfobs, sam, sepa_evo, eccen_evo, nharms=100, nreals=None, *, seed=None,
                                       harm_chunk=None):
    """(F, nharms, nreals) -- or (F, nharms) for nreals=None -- Poisson-sampled GWB strain squared
    per harmonic of an eccentric SAM (gravwaves.py:944-981)."""
    from .cyutils import _seed
    squeeze = nreals is None
    R = 1 if squeeze else int(nreals)
    gwb = D.to_host(engine.gwb_eccen_discrete(*_eccen_inputs(gwfobs, sam, sepa_evo, eccen_evo), int(nharms), R,
                                              seed=_seed(seed), harm_chunk=harm_chunk))
    return gwb.squeeze() if squeeze else gwb


def _gwb_ideal_device(fobs_gw, dens, log10_mtot_d, mtot_d, mrat_d, redz_d, sum=True):
    """Device evaluation of the idealised GWB (gravwaves.py:619-663); dens (M,Q,Z) [Mpc^-3],
    redz_d (Z) or (M,Q,Z)."""
    from .constants import NWTG, SPLC, MPC
    M, Q, Z = dens.shape
    fo = D.to_dev(np.atleast_1d(np.asarray(fobs_gw, dtype=float)))
    F = fo.numel()
    rz3 = redz_d.dim() == 3
    if sum:
        ws = D.workspace(_lib.size("hd_gwb_ideal_workspace_size"))
        out = D.empty(F)
    else:
        ws = None
        out = D.empty((M - 1, Q - 1, Z - 1, F))
    _lib.call("hd_gwb_ideal", D.ptr(dens), D.ptr(log10_mtot_d), D.ptr(mtot_d), M, D.ptr(mrat_d), Q,
              D.ptr(redz_d.contiguous()), int(rz3), Z, D.ptr(fo), F, float(NWTG), float(SPLC), float(MPC), int(bool(sum)),
              D.ptr(out), D.ptr(ws), 0 if ws is None else ws.numel(), D.stream_ptr())
    res = D.to_host(out)
    return res if np.ndim(fobs_gw) else res.reshape(res.shape[:-1])


def gwb_ideal(fobs_gw, ndens, mtot, mrat, redz, dlog10, sum=True):
    """Idealised GWB h_c (gravwaves.py:619-663): ``ndens`` (M,Q,Z) [cm^-3] at the grid nodes,
    ``mtot`` (M,1,1) / ``mrat`` (1,Q,1) edges, ``redz`` (1,1,Z) edges or (M,Q,Z) node redshifts."""
    from .constants import MPC
    if not dlog10:
        raise NotImplementedError("only the SAM's integration in log10 of the total mass is provided")
    ndens = np.asarray(ndens, dtype=float)
    mtot = np.ravel(mtot)
    mrat = np.ravel(mrat)
    redz = np.asarray(redz, dtype=float)
    rz = redz if redz.shape == ndens.shape else np.ravel(redz)
    return _gwb_ideal_device(fobs_gw, D.to_dev(ndens * MPC ** 3), D.to_dev(np.log10(mtot)), D.to_dev(mtot),
                             D.to_dev(mrat), D.to_dev(rz), sum=sum)
