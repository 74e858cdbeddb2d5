"""Binary hardening models of the SAM hot path (holodeck/hardening.py:87-212, 1371-1452).

``Hard_GW``: GW-only inspiral.  ``Fixed_Time_2PL_SAM``: phenomenological
double power law plus GW, normalised per (M, q) so that the lifetime from
``sepa_init`` to the ISCO equals ``time``; the normalisations are found on the
device by Brent's method (hd_find_2pwl_norm).
"""
import numpy as np

from . import _lib
from . import device as D
from .constants import PC, GYR


class _Hardening:
    CONSISTENT = None

    def __str__(self):
        return f"{self.__class__.__name__}"


class Hard_GW(_Hardening):
    """Gravitational-wave driven hardening (hardening.py:87-212)."""
    CONSISTENT = False

    @property
    def consistent(self):
        return False


class Fixed_Time_2PL_SAM(_Hardening):
    """SAM-optimised fixed-lifetime double power law (hardening.py:1371-1452)."""
    CONSISTENT = True

    def __init__(self, sam, time, sepa_init=1.0e3 * PC, rchar=10.0 * PC, gamma_inner=-1.0, gamma_outer=+1.5,
                 num_steps=300):
        assert np.ndim(time) == 0
        assert np.ndim(rchar) == 0
        assert np.ndim(gamma_inner) == 0
        assert np.ndim(gamma_outer) == 0
        from . import engine
        mtot, mrat = np.meshgrid(sam.mtot, sam.mrat, indexing='ij')
        shape = mtot.shape
        norm_log10, norm = engine.find_2pwl_norm(time, mtot.ravel(), mrat.ravel(), sepa_init, rchar, gamma_inner,
                                                 gamma_outer, num_steps)
        self._norm_log10_dev = norm_log10.reshape(shape)
        self._norm_dev = norm.reshape(shape)
        self._target_time = time
        self._num_steps = num_steps
        self._sepa_init = sepa_init
        self._rchar = rchar
        self._gamma_inner = gamma_inner
        self._gamma_outer = gamma_outer

    @property
    def _norm(self):
        """(M, Q) hardening normalisations 10**norm_log10 (host copy)."""
        return D.to_host(self._norm_dev)

    def __str__(self):
        return (f"{super().__str__()} :: target_time/Gyr={self._target_time / GYR:.2e} num_steps={self._num_steps} "
                f"sepa_init/pc={self._sepa_init / PC:.2e} rchar/pc={self._rchar / PC:.2e} "
                f"gamma_inner={self._gamma_inner:.2e} gamma_outer={self._gamma_outer:.2e} ")

    def dadt(self, mtot, mrat, sepa, norm=None):
        """Hardening rate [cm/s] (negative), evaluated on the device (sam_cyutils.pyx:271-286)."""
        if norm is None:
            norm = self._norm
        args = np.broadcast_arrays(mtot, mrat, sepa, norm)
        shape = args[0].shape
        mt, mr, sa, nm = [D.to_dev(np.ascontiguousarray(a, dtype=float).ravel()) for a in args]
        out = D.empty(mt.numel())
        _lib.call("hd_hard_func_2pwl_gw", D.ptr(mt), D.ptr(mr), D.ptr(sa), D.ptr(nm), mt.numel(), float(self._rchar),
                  float(self._gamma_inner), float(self._gamma_outer), D.ptr(out), D.stream_ptr())
        return D.to_host(out).reshape(shape)

