"""Oracle of the idealised GWB: gravwaves.gwb_ideal (gravwaves.py:619-663), restated in numpy
with the same successive trapezoid steps over log10 M, q and z, including the reference's
down-averaging of a per-node redshift array along the already-integrated axes.

TEST INFRASTRUCTURE ONLY.
"""
import numpy as np

NWTG = 6.6743e-08               # astropy G [cgs] (holodeck/constants.py)
SPLC = 29979245800.0            # c [cm/s]
MPC = 3.085677581491367e+24     # Mpc [cm]


def chirp_mass_mtmr(mt, mr):
    """utils.chirp_mass_mtmr (utils.py:1978-1997)."""
    return mt * np.power(mr, 3.0 / 5.0) / np.power(1 + mr, 6.0 / 5.0)


def gwb_ideal(fobs_gw, ndens, mtot, mrat, redz, dlog10=True, sum=True):
    """ndens (M,Q,Z) [cm^-3]; mtot (M,1,1), mrat (1,Q,1), redz (1,1,Z) or (M,Q,Z)."""
    const = (4.0 * np.pi) / (3 * SPLC ** 2)
    mc = np.power(NWTG * chirp_mass_mtmr(mtot, mrat), 5.0 / 3.0)
    with np.errstate(divide='ignore'):
        rz = np.power(1 + redz, -1.0 / 3.0)     # inf at the z = -1 sentinel, zeroed below
    fogw = np.power(np.pi * np.asarray(fobs_gw, dtype=float), -4.0 / 3.0)
    with np.errstate(invalid='ignore'):
        integ = ndens * mc * rz                 # 0 * inf at empty sentinel nodes, zeroed below
    redz = redz * np.ones_like(integ)
    integ = np.where(redz <= 0.0, 0.0, integ)
    args = [np.log10(mtot) if dlog10 else mtot, mrat, redz]
    for ax, xx in enumerate(args):
        integ = np.moveaxis(integ, ax, 0)
        xx = np.moveaxis(xx, ax, 0)
        if xx.shape[1:] != (1,) * (xx.ndim - 1):
            # per-node coordinates: average them down along the axes already integrated
            for jj in range(1, xx.ndim):
                if xx.shape[jj] not in (1, integ.shape[jj]):
                    xx = np.moveaxis(xx, jj, 0)
                    xx = 0.5 * (xx[:-1] + xx[1:])
                    xx = np.moveaxis(xx, 0, jj)
        integ = 0.5 * (integ[:-1] + integ[1:]) * np.diff(xx, axis=0)
        integ = np.moveaxis(integ, 0, ax)
    gwb = const * fogw
    gwb = gwb * np.sum(integ) if sum else gwb * integ[..., np.newaxis]
    return np.sqrt(gwb)
