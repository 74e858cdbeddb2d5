"""Drop-in for the realization kernels of ``holodeck.cyutils`` (holodeck/cyutils.pyx:854-1767).

Same names, argument order and return tuples; numpy in and out, device
compute.  Keyword-only extensions: ``seed`` (the reference's PCG64 is
unseeded) and ``counts`` (inject the per-(bin, frequency, realization) counts,
shape number.shape + (R,), instead of drawing them).
"""
import numpy as np

from . import _lib
from . import device as D
from . import engine


def _prep(number, h2fdf):
    number = np.asarray(number, dtype=float)
    F = number.shape[-1]
    n = D.to_dev(number).reshape(-1, F).contiguous()
    h = D.to_dev(np.asarray(h2fdf, dtype=float)).reshape(-1, F).contiguous()
    return n, h, number.shape


def _seed(seed):
    if seed is None:
        seed = int(np.random.SeedSequence().generate_state(1, dtype=np.uint64)[0])
    return int(seed)


def _counts(counts, B, F):
    return None if counts is None else D.to_dev(np.asarray(counts, dtype=float)).reshape(B, F, -1).contiguous()


def _long_threshold(t):
    """The reference's Cython loops take the threshold as ``long thresh`` (truncated toward zero)."""
    t = float(t)
    return float(int(t)) if np.isfinite(t) else t


def _run(number, h2fdf, nreals, nloudest, mode, order=None, flags=0, par=None, normal_threshold=1e10, seed=None,
         counts=None):
    normal_threshold = _long_threshold(normal_threshold)
    n, h, shape = _prep(number, h2fdf)
    B, F = h.shape
    cnt = _counts(counts, B, F)
    R = cnt.shape[-1] if cnt is not None else int(nreals)
    return engine.realize(n, h, R, nloudest, mode, order=order, counts=cnt, flags=flags, par=par,
                          seed=_seed(seed), normal_threshold=normal_threshold), shape


def _order_from_sorts(shape3, msort, qsort, zsort):
    flat = np.ravel_multi_index((np.asarray(msort), np.asarray(qsort), np.asarray(zsort)), shape3)
    return D.to_dev(flat.astype(np.int64))


def loudest_hc_from_sorted(number, h2fdf, nreals, nloudest, msort, qsort, zsort, normal_threshold=1e10, *,
                           seed=None, counts=None):
    """-> (hc2ss (F,R,L), hc2bg (F,R)) (cyutils.pyx:1220-1260)."""
    shape3 = np.shape(number)[:3]
    order = _order_from_sorts(shape3, msort, qsort, zsort)
    (hc2ss, _, hc2bg, _), _ = _run(number, h2fdf, nreals, nloudest, _lib.HD_SEL_SORTED, order=order,
                                   normal_threshold=normal_threshold, seed=seed, counts=counts)
    return D.to_host(hc2ss), D.to_host(hc2bg)


def loudest_hc_and_par_from_sorted(number, h2fdf, nreals, nloudest, mt, mr, rz, msort, qsort, zsort,
                                   normal_threshold=1e10, *, seed=None, counts=None):
    """-> (hc2ss (F,R,L), hc2bg (F,R), lspar (3,F,R), bgpar (3,F,R), ssidx (3,F,R,L))
    (cyutils.pyx:1347-1538).

    The frequency-0 order walk with the strain-weighted mean (M, q, z) of the loud instances
    (lspar; 0/0 = NaN where there is none, as the reference's C division) and of the background;
    zero-strain bins are not skipped (unlike the _redz variant).  ssidx holds the (M, Q, Z) bin
    indices of each loud slot, 0 where a slot is empty (the reference's np.zeros).  The threshold
    is taken as an integer, as the reference's ``long thresh`` argument."""
    number = np.asarray(number, dtype=float)
    shape3 = number.shape[:3]
    F = number.shape[-1]
    mt, mr, rz = (np.asarray(a, dtype=float) for a in (mt, mr, rz))
    mm, qq, zz = np.meshgrid(np.arange(shape3[0]), np.arange(shape3[1]), np.arange(shape3[2]), indexing='ij')
    pars = [mt[mm], mr[qq], rz[zz]]
    par = D.to_dev(np.stack([np.repeat(p.reshape(-1, 1), F, axis=1) for p in pars]))
    order = _order_from_sorts(shape3, msort, qsort, zsort)
    (hc2ss, ssidx, hc2bg, bgpar), _ = _run(number, h2fdf, nreals, nloudest, _lib.HD_SEL_SORTED, order=order,
                                           par=par, normal_threshold=normal_threshold, seed=seed, counts=counts)
    ss = D.to_host(hc2ss)
    idx = D.to_host(ssidx)
    ok = idx >= 0
    m3 = np.unravel_index(np.where(ok, idx, 0), shape3)
    out_idx = np.zeros((3,) + idx.shape, dtype=np.int64)
    for k in range(3):
        out_idx[k][ok] = m3[k][ok]
    vals = (mt[m3[0]], mr[m3[1]], rz[m3[2]])
    sum_ls = ss.sum(axis=-1)
    with np.errstate(invalid='ignore', divide='ignore'):
        lspar = np.stack([np.where(ok, ss * v, 0.0).sum(axis=-1) / sum_ls for v in vals])
    return ss, D.to_host(hc2bg), lspar, D.to_host(bgpar), out_idx


def loudest_hc_and_par_from_sorted_redz(number, h2fdf, nreals, nloudest, mt, mr, rz, redz_final, dcom_final, sepa,
                                        angs, msort, qsort, zsort, normal_threshold=1e10, *, seed=None, counts=None):
    """-> (hc2ss (F,R,L), hc2bg (F,R), sspar (4,F,R,L), bgpar (7,F,R)) (cyutils.pyx:1541-1608)."""
    number = np.asarray(number, dtype=float)
    shape3 = number.shape[:3]
    F = number.shape[-1]
    mm, qq, zz = np.meshgrid(np.arange(shape3[0]), np.arange(shape3[1]), np.arange(shape3[2]), indexing='ij')
    full = lambda a: np.broadcast_to(np.asarray(a, float)[..., None], shape3 + (F,))   # noqa: E731
    pars = [full(np.asarray(mt)[mm]), full(np.asarray(mr)[qq]), full(np.asarray(rz)[zz]), redz_final, dcom_final,
            sepa, angs]
    par = D.to_dev(np.stack([np.broadcast_to(np.asarray(p, float), shape3 + (F,)).reshape(-1, F) for p in pars]))
    order = _order_from_sorts(shape3, msort, qsort, zsort)
    (hc2ss, ssidx, hc2bg, bgpar), _ = _run(number, h2fdf, nreals, nloudest, _lib.HD_SEL_SORTED, order=order,
                                           flags=_lib.HD_FLAG_SKIP_ZERO_H, par=par,
                                           normal_threshold=normal_threshold, seed=seed, counts=counts)
    idx = D.to_host(ssidx)
    ok = idx >= 0
    sspar = np.zeros((4,) + idx.shape)
    m3 = np.unravel_index(np.where(ok, idx, 0), shape3)
    fidx = np.broadcast_to(np.arange(F)[:, None, None], idx.shape)
    zf = np.asarray(redz_final, float)
    vals = [np.asarray(mt)[m3[0]], np.asarray(mr)[m3[1]], np.asarray(rz)[m3[2]], zf[m3[0], m3[1], m3[2], fidx]]
    for k in range(4):
        sspar[k][ok] = vals[k][ok]
    return D.to_host(hc2ss), D.to_host(hc2bg), sspar, D.to_host(bgpar)


def ss_bg_hc(number, h2fdf, nreals, normal_threshold=1e10, *, seed=None, counts=None):
    """-> (hc2ss (F,R), hc2bg (F,R), ssidx (3,F,R)) (cyutils.pyx:900-932)."""
    shape3 = np.shape(number)[:3]
    (hc2ss, ssidx, hc2bg, _), _ = _run(number, h2fdf, nreals, 1, _lib.HD_SEL_ARGMAX,
                                       normal_threshold=normal_threshold, seed=seed, counts=counts)
    idx = D.to_host(ssidx)[..., 0]
    out = np.full((3,) + idx.shape, -1, dtype=np.int64)
    ok = idx >= 0
    m3 = np.unravel_index(idx[ok], shape3)
    for k in range(3):
        out[k][ok] = m3[k]
    return D.to_host(hc2ss)[..., 0], D.to_host(hc2bg), out


def ss_bg_hc_and_par(number, h2fdf, nreals, mt, mr, rz, normal_threshold=1e10, *, seed=None, counts=None):
    """-> (hc2ss, hc2bg, ssidx (3,F,R), bgpar (3,F,R), sspar (3,F,R)) (cyutils.pyx:1017-1062)."""
    number = np.asarray(number, dtype=float)
    shape3 = number.shape[:3]
    F = number.shape[-1]
    mm, qq, zz = np.meshgrid(np.arange(shape3[0]), np.arange(shape3[1]), np.arange(shape3[2]), indexing='ij')
    pars = [np.asarray(mt, float)[mm], np.asarray(mr, float)[qq], np.asarray(rz, float)[zz]]
    par = D.to_dev(np.stack([np.repeat(p.reshape(-1, 1), F, axis=1) for p in pars]))
    (hc2ss, ssidx, hc2bg, bgpar), _ = _run(number, h2fdf, nreals, 1, _lib.HD_SEL_ARGMAX, par=par,
                                           normal_threshold=normal_threshold, seed=seed, counts=counts)
    idx = D.to_host(ssidx)[..., 0]
    ok = idx >= 0
    out = np.full((3,) + idx.shape, -1, dtype=np.int64)
    m3 = np.unravel_index(idx[ok], shape3)
    for k in range(3):
        out[k][ok] = m3[k]
    sspar = np.zeros((3,) + idx.shape)
    sspar[0][ok] = np.asarray(mt)[m3[0]]
    sspar[1][ok] = np.asarray(mr)[m3[1]]
    sspar[2][ok] = np.asarray(rz)[m3[2]]
    return D.to_host(hc2ss)[..., 0], D.to_host(hc2bg), out, D.to_host(bgpar), sspar


def sam_poisson_gwb(dist, hc2, nreals, normal_threshold=1e10, *, seed=None, counts=None):
    """sum_b Poisson(dist) * hc2 -> (F, R) (cyutils.pyx:854-897)."""
    (_, _, hc2bg, _), _ = _run(dist, hc2, nreals, 0, _lib.HD_SEL_NONE, normal_threshold=normal_threshold, seed=seed,
                               counts=counts)
    return D.to_host(hc2bg)


def _eccen_dev(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo):
    return (D.to_dev(np.asarray(ndens, dtype=float)).contiguous(), D.to_dev(mtot_log10), D.to_dev(mrat),
            D.to_dev(redz), D.to_dev(dcom), D.to_dev(np.atleast_1d(gwfobs)), D.to_dev(sepa_evo), D.to_dev(eccen_evo))


def sam_calc_gwb_single_eccen(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms=100):
    """-> gwb (F, nharms): characteristic strain squared per frequency and harmonic
    (cyutils.pyx:361-597).  ndens (M,Q,Z) at the grid nodes [Mpc^-3], dcom [Mpc]."""
    args = _eccen_dev(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo)
    return D.to_host(engine.gwb_eccen(*args, int(nharms)))


def sam_calc_gwb_single_eccen_discrete(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo, nharms,
                                       nreals, *, seed=None, harm_chunk=None):
    """-> gwb (F, nharms, nreals) with Poisson counts per (node, frequency, harmonic, realization)
    (cyutils.pyx:600-851)."""
    args = _eccen_dev(ndens, mtot_log10, mrat, redz, dcom, gwfobs, sepa_evo, eccen_evo)
    return D.to_host(engine.gwb_eccen_discrete(*args, int(nharms), int(nreals), seed=_seed(seed),
                                               harm_chunk=harm_chunk))
