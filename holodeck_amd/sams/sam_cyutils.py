"""Drop-in for ``holodeck.sams.sam_cyutils`` (holodeck/sams/sam_cyutils.pyx).

Same function names, argument order and return tuples; numpy FP64 in and out.
The work runs on the device (libholohip.so).

``dynamic_binary_number_at_fobs(fobs_orb, sam, hard, cosmo)`` reads exactly what the reference
reads (sam_cyutils.pyx:457-494) from whatever objects it is handed -- this package's classes or the
reference's own ``Semi_Analytic_Model``, ``Fixed_Time_2PL_SAM`` / ``Hard_GW`` and cosmopy
``Cosmology`` -- so the reference can swap its Cython module for this one:

* ``sam.static_binary_density`` (M,Q,Z), ``sam.shape``, ``sam.mtot``, ``sam.mrat``, ``sam.redz``,
  ``sam._gmt_time`` (2PL; None -> zeros), ``sam._redz_prime`` (GW; None -> the redshift edges);
* ``hard._sepa_init``, ``_num_steps``, ``_norm`` (M,Q), ``_rchar``, ``_gamma_inner``,
  ``_gamma_outer``;
* ``cosmo._grid_z`` (decreasing, to z=0), ``_grid_dcom`` [cm], ``_grid_age`` [s].

The hardening model is dispatched as the reference does, by class (``Fixed_Time_2PL_SAM`` or
``Hard_GW``, an instance or the class itself for GW), matched by this package's classes or by name
in the object's class hierarchy.
"""
import numpy as np

from .. import _lib
from .. import device as D
from .. import engine


def _class_names(obj):
    cls = obj if isinstance(obj, type) else type(obj)
    return {c.__name__ for c in cls.__mro__}


def _is_2pwl(hard):
    from .. import hardening
    return isinstance(hard, hardening.Fixed_Time_2PL_SAM) or "Fixed_Time_2PL_SAM" in _class_names(hard)


def _is_gw(hard):
    from .. import hardening
    if isinstance(hard, hardening.Hard_GW) or (isinstance(hard, type) and issubclass(hard, hardening.Hard_GW)):
        return True
    return "Hard_GW" in _class_names(hard)


def _dev(a):
    return D.to_dev(np.ascontiguousarray(a, dtype=np.float64))


def dynamic_binary_number_at_fobs(fobs_orb, sam, hard, cosmo):
    """-> (redz_final, diff_num), each (M, Q, Z, F) (sam_cyutils.pyx:421-504)."""
    fobs_orb = np.atleast_1d(np.asarray(fobs_orb, dtype=float))
    F = fobs_orb.size
    nden = np.asarray(sam.static_binary_density, dtype=float)
    shape = tuple(sam.shape)
    if nden.shape != shape:
        raise ValueError(f"static_binary_density shape {nden.shape} != sam.shape {shape}")
    M, Q, Z = shape
    gz = _dev(cosmo._grid_z)
    gd = _dev(cosmo._grid_dcom)
    if gz.numel() != gd.numel() or gz.numel() < 2:
        raise ValueError("cosmo._grid_z and _grid_dcom must have the same length >= 2")
    fo = _dev(fobs_orb)
    nd = _dev(nden)
    mt, mr, rz = _dev(sam.mtot), _dev(sam.mrat), _dev(sam.redz)
    rf = D.empty(shape + (F,))
    dn = D.empty(shape + (F,))
    if _is_2pwl(hard):
        gmt_time = sam._gmt_time
        if gmt_time is None:          # a SAM with a merger rate (GMR) has no merger times
            gmt_time = np.zeros(shape)
        ga = _dev(cosmo._grid_age)
        if ga.numel() != gz.numel():
            raise ValueError("cosmo._grid_age must have the length of _grid_z")
        norm = _dev(hard._norm)
        if norm.numel() != M * Q:
            raise ValueError(f"hard._norm has {norm.numel()} values, the grid needs {M * Q}")
        srt = bool(np.all(np.diff(fobs_orb) >= 0))
        _lib.call("hd_dbn_2pwl", D.ptr(fo), F, int(srt), float(hard._sepa_init), int(hard._num_steps), D.ptr(norm),
                  float(hard._rchar), float(hard._gamma_inner), float(hard._gamma_outer), D.ptr(nd), D.ptr(mt), M,
                  D.ptr(mr), Q, D.ptr(rz), Z, D.ptr(_dev(gmt_time)), D.ptr(gz), D.ptr(gd), D.ptr(ga), gz.numel(),
                  D.ptr(rf), D.ptr(dn), D.stream_ptr())
    elif _is_gw(hard):
        redz_prime = sam._redz_prime
        if redz_prime is None:        # no merger delay: the initial redshifts
            redz_prime = np.asarray(sam.redz, dtype=float)[np.newaxis, np.newaxis, :] * np.ones(shape)
        _lib.call("hd_dbn_gw", D.ptr(fo), F, D.ptr(nd), D.ptr(mt), M, D.ptr(mr), Q, Z, D.ptr(_dev(redz_prime)),
                  D.ptr(gz), D.ptr(gd), gz.numel(), D.ptr(rf), D.ptr(dn), D.stream_ptr())
    else:
        raise ValueError(f"Unexpected `hard` value {hard}!")
    return D.to_host(rf), D.to_host(dn)


def integrate_differential_number_3dx1d(edges, dnum):
    """(M,Q,Z,F) differential number -> (M-1,Q-1,Z-1,F) numbers (sam_cyutils.pyx:115-163)."""
    shape = [len(ee) for ee in edges]
    shape[-1] -= 1
    assert np.shape(dnum) == tuple(shape), f"Shape of edges={shape} does not match dnum={np.shape(dnum)}"
    lm = D.to_dev(np.log10(edges[0]))
    mr = D.to_dev(np.asarray(edges[1], dtype=float))
    rz = D.to_dev(np.asarray(edges[2], dtype=float))
    dlnf = D.to_dev(np.diff(np.log(edges[3])))
    out = engine.integrate_3dx1d(lm, mr, rz, dlnf, D.to_dev(dnum))
    return D.to_host(out)


def find_2pwl_hardening_norm(time, mtot, mrat, sepa_init, rchar, gamma_inner, gamma_outer, nsteps):
    """log10 normalisations (sam_cyutils.pyx:289-306)."""
    assert np.ndim(time) == 0
    assert np.ndim(mtot) == 1
    assert np.shape(mtot) == np.shape(mrat)
    out, _ = engine.find_2pwl_norm(time, np.asarray(mtot, float), np.asarray(mrat, float), sepa_init, rchar,
                                   gamma_inner, gamma_outer, nsteps)
    return D.to_host(out)


def integrate_binary_evolution_2pwl(norm_log10, mtot, mrat, sepa_init, rchar, gamma_inner, gamma_outer, nsteps):
    """Lifetime [s] from ``sepa_init`` to the ISCO at a given normalisation (sam_cyutils.pyx:401-418).

    The reference takes scalars; arrays broadcast and return that shape."""
    args = np.broadcast_arrays(norm_log10, mtot, mrat)
    shape = args[0].shape
    nl, mt, mr = [_dev(np.ravel(a)) for a in args]
    out = D.empty(mt.numel())
    _lib.call("hd_lifetime_2pwl", D.ptr(nl), D.ptr(mt), D.ptr(mr), mt.numel(), float(sepa_init), float(rchar),
              float(gamma_inner), float(gamma_outer), int(nsteps), D.ptr(out), D.stream_ptr())
    res = D.to_host(out).reshape(shape)
    return float(res) if res.ndim == 0 else res


def hard_func_2pwl_gw(mtot, mrat, sepa, norm, rchar, gamma_inner, gamma_outer):
    """2PL + GW hardening rate, broadcast over the array arguments (sam_cyutils.pyx:271-286)."""
    args = np.broadcast_arrays(mtot, mrat, sepa, norm)
    shape = args[0].shape
    mt, mr, sa, nm = [D.to_dev(np.ascontiguousarray(a, dtype=float).ravel()) for a in args]
    out = D.empty(mt.numel())
    _lib.call("hd_hard_func_2pwl_gw", D.ptr(mt), D.ptr(mr), D.ptr(sa), D.ptr(nm), mt.numel(), float(rchar),
              float(gamma_inner), float(gamma_outer), D.ptr(out), D.stream_ptr())
    return D.to_host(out).reshape(shape)
