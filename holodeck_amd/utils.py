"""Small host-side helpers over 1-D grids (frequency bins, bin edges, integer checks).

Only O(F), O(M), O(Q), O(Z) bookkeeping lives here; every per-cell quantity of
the hot path is computed on the device.
"""
import numbers

import numpy as np

from .constants import YR


def pta_freqs(dur=16.03 * YR, num=40, cad=None):
    """Fourier frequency bins of a PTA of duration `dur` (utils.py:835-874).

    Returns (cents, edges): centres i/dur for i=1..num and edges (i-1/2)/dur.
    """
    fmin = 1.0 / dur
    if cad is not None:
        num = int(np.floor(dur / (2.0 * cad)))
    cents = np.arange(1, num + 2) * fmin
    edges = cents - fmin / 2.0
    cents = cents[:-1]
    return cents, edges


def midpoints(vals, axis=-1, log=False):
    """Arithmetic (or log10) midpoints along `axis` (utils.py:703-711)."""
    mm = np.moveaxis(np.asarray(vals), axis, 0)
    if log:
        mm = np.log10(mm)
    mm = 0.5 * (mm[1:] + mm[:-1])
    if log:
        mm = 10.0 ** mm
    return np.moveaxis(mm, 0, axis)


def isnumeric(val):
    try:
        float(str(val))
    except ValueError:
        return False
    return True


def isinteger(val):
    """Integral number test (utils.py:656-671)."""
    return isnumeric(val) and isinstance(val, numbers.Integral)


def m1m2_from_mtmr(mt, mr):
    """(M, q) -> (m1, m2) with m1 the primary (utils.py:1620-1642)."""
    mt = np.asarray(mt)
    mr = np.asarray(mr)
    m1 = mt / (1.0 + mr)
    m2 = mt - m1
    return np.array([m1, m2])


def chirp_mass_mtmr(mt, mr):
    """Chirp mass from total mass and ratio (utils.py:1978-1997)."""
    mt = np.asarray(mt)
    mr = np.asarray(mr)
    return mt * np.power(mr, 3.0 / 5.0) / np.power(1 + mr, 6.0 / 5.0)


def _parse_val_log10_val_pars(val, val_log10, val_units=1.0, name='value', only_one=True):
    """Either a value or its log10 (utils.py:1299-1337)."""
    both_or_neither = ((val_log10 is not None) == (val is not None))
    if only_one and both_or_neither:
        raise ValueError(f"One of {name} OR {name}_log10 must be provided!  {name}={val}, {name}_log10={val_log10}")
    if val is None:
        val = val_units * np.power(10.0, val_log10)
    if val_log10 is None:
        val_log10 = np.log10(val / val_units)
    return val, val_log10


def get_subclass_instance(value, default, superclass, allow_none=False):
    """Instantiate `value` (class or instance) of `superclass`, or `default` (utils.py:320-367)."""
    import inspect
    if value is None:
        value = default
    if value is None:
        if allow_none:
            return None
        raise ValueError(f"No value given for {superclass}!")
    if inspect.isclass(value):
        value = value()
    if not isinstance(value, superclass):
        raise ValueError(f"argument ({value}) is not a subclass of `{superclass}`!")
    return value
