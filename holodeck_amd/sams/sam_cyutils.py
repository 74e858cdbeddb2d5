"""Drop-in for ``holodeck.sams.sam_cyutils`` (holodeck/sams/sam_cyutils.pyx).

Same function names, argument order and return tuples; numpy FP64 in and out.
The work runs on the device (libholohip.so).
"""
import numpy as np

from .. import _lib
from .. import device as D
from .. import engine


def dynamic_binary_number_at_fobs(fobs_orb, sam, hard, cosmo=None):
    """-> (redz_final, diff_num), each (M, Q, Z, F) (sam_cyutils.pyx:421-504)."""
    rf, dn = sam._dbn_device(hard, fobs_orb)
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


def hard_func_2pwl_gw(mtot, mrat, sepa, norm, rchar, gamma_inner, gamma_outer):
    """2PL + GW hardening rate, broadcast over the array arguments (sam_cyutils.pyx:271-286)."""
    args = np.broadcast_arrays(mtot, mrat, sepa, norm)
    shape = args[0].shape
    mt, mr, sa, nm = [D.to_dev(np.ascontiguousarray(a, dtype=float).ravel()) for a in args]
    out = D.empty(mt.numel())
    _lib.call("hd_hard_func_2pwl_gw", D.ptr(mt), D.ptr(mr), D.ptr(sa), D.ptr(nm), mt.numel(), float(rchar),
              float(gamma_inner), float(gamma_outer), D.ptr(out), D.stream_ptr())
    return D.to_host(out).reshape(shape)
