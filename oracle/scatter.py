"""Oracle of the M-Mbulge scatter step: add_scatter_to_masses (holodeck/sams/sam.py:1291-1394)
restated with the same scipy routines the reference calls (CloughTocher2DInterpolator,
NearestNDInterpolator, RegularGridInterpolator, scipy.stats.norm) and the same
weight construction and einsum passes (holodeck/utils.py:382-488).

TEST INFRASTRUCTURE ONLY.
"""
import numpy as np
import scipy as sp
import scipy.interpolate  # noqa
import scipy.stats  # noqa


def _roll_rows(arr, roll_num):
    """utils.roll_rows (utils.py:382-413)."""
    roll = np.asarray(roll_num)
    nrows, ncols = arr.shape
    arr_roll = arr[:, [*range(ncols), *range(ncols - 1)]].copy()
    strd_0, strd_1 = arr_roll.strides
    result = np.lib.stride_tricks.as_strided(arr_roll, (nrows, ncols, ncols), (strd_0, strd_1, strd_1))
    return result[np.arange(nrows), (ncols - roll) % ncols]


def _get_scatter_weights(uniform_cents, dist):
    """utils.get_scatter_weights (utils.py:416-452)."""
    num = uniform_cents.size
    dx = np.diff(uniform_cents)
    dx = dx[0]
    dx = dx / 2.0 + np.arange(num) * dx
    dx = np.concatenate([-dx[::-1], dx])
    return np.diff(dist.cdf(dx))


def _get_rolled_weights(log_cents, dist):
    """utils._get_rolled_weights (utils.py:464-488)."""
    num = log_cents.size
    weights = _get_scatter_weights(log_cents, dist)
    weights = weights[np.newaxis, :] * np.ones((num, weights.size))
    roll = 1 - num + np.arange(num)
    weights = _roll_rows(weights, roll)
    return weights[:, :num]


def _scatter_with_weights(dens, weights, axis=0):
    """utils._scatter_with_weights (utils.py:455-461), including its output axis order."""
    dens = np.moveaxis(dens, axis, 0)
    dens_new = np.einsum("j...,jk...", dens, weights)
    return np.moveaxis(dens_new, 0, axis)


def add_scatter_to_masses(mtot, mrat, dens, scatter, refine=4, return_grads=False):
    dist = sp.stats.norm(loc=0.0, scale=scatter)
    output = np.zeros_like(dens)
    m1 = mtot[:, np.newaxis] / (1.0 + mrat[np.newaxis, :])
    m2 = mtot[:, np.newaxis] - m1
    m1m2_on_mtmr_grid = (m1.flatten(), m2.flatten())
    grid_size = m1.shape[0] * refine
    ext = [0.9 * mtot[0] * mrat[0] / (1.0 + mrat[0]), mtot[-1] * (1.0 + mrat[0]) / mrat[0]]
    mextr = np.array([np.min(ext), np.max(ext)])
    _mgrid = np.logspace(*np.log10(mextr), grid_size)
    mgrid_log10 = np.log10(_mgrid)
    m1m2_on_mtmr_grid = tuple([np.log10(mm) for mm in m1m2_on_mtmr_grid])
    m1m2_grid = np.meshgrid(mgrid_log10, mgrid_log10, indexing='ij')
    numz = np.shape(dens)[2]
    dlay = None
    weights = _get_rolled_weights(mgrid_log10, dist)
    grads = []
    for ii in range(numz):
        dens_redz = dens[:, :, ii]
        points = m1m2_on_mtmr_grid if dlay is None else dlay
        interp = sp.interpolate.CloughTocher2DInterpolator(points, dens_redz.flatten())
        m1m2_dens = interp(tuple(m1m2_grid))
        if return_grads:
            grads.append(np.array(interp.grad).reshape(-1, 2))
        if dlay is None:
            dlay = interp.tri
        bads = np.isnan(m1m2_dens) | (m1m2_dens < 0.0)
        if np.any(bads):
            interp = sp.interpolate.NearestNDInterpolator(points, dens_redz.flatten())
            temp = interp(tuple(m1m2_grid))
            m1m2_dens[bads] = temp[bads]
            bads = np.isnan(m1m2_dens) | (m1m2_dens < 0.0)
            if np.any(bads):
                raise ValueError("After 0th order interpolation, bad values remain!")
        m1m2_dens = _scatter_with_weights(m1m2_dens, weights, axis=0)
        m1m2_dens = _scatter_with_weights(m1m2_dens, weights, axis=1)
        interp = sp.interpolate.RegularGridInterpolator((mgrid_log10, mgrid_log10), m1m2_dens)
        m1m2_dens = interp(m1m2_on_mtmr_grid, method='linear').reshape(m1.shape)
        output[:, :, ii] = m1m2_dens[...]
    if return_grads:
        return output, grads
    return output
