"""M-Mbulge scatter of the binary density on the device.

<- add_scatter_to_masses (holodeck/sams/sam.py:1291-1394) with
   utils._get_rolled_weights / get_scatter_weights / _scatter_with_weights
   (holodeck/utils.py:416-488).

Per redshift slice the reference (i) Clough-Tocher-interpolates the density
from the (log10 m1, log10 m2) nodes of the (M, q) grid onto a regular
(4M)x(4M) log-mass grid, (ii) replaces NaN / negative values with the nearest
node's value, (iii) convolves with Gaussian-CDF weights, (iv) interpolates
bilinearly back to the (M, q) nodes.

Note on (iii): ``_scatter_with_weights`` returns its einsum result with the
convolved axis moved last, so the reference's two calls (axis=0, then axis=1)
both act on the m1 axis; the slice result is (W W)^T D (D indexed [m1, m2]).
This is reproduced as is.

Split of work:
* the PLAN -- geometry only, built once per (mtot, mrat) on the host:
  the Delaunay triangulation of the nodes (scipy.spatial.Delaunay / Qhull, as
  the reference's interpolator builds it), vertex adjacency in Qhull's order,
  a level schedule of the Gauss-Seidel sweep, the simplex and barycentric
  coordinates of every regular-grid point, each simplex's cross-boundary
  derivative directions, the nearest node of every regular-grid point and the
  bilinear weights back to the nodes; the Gaussian-CDF weights of a scatter
  value are made per call (P = (W W)^T by an FP64 GEMM on the device);
* the DEVICE -- everything that depends on the density: the global gradient
  estimation (Gauss-Seidel, level-scheduled, identical update sequence), the
  Clough-Tocher evaluation and bad-value fill, the FP64 GEMM with (W W)^T and
  the bilinear gather, for all redshift slices at once.
"""
import atexit
import collections

import numpy as np
import torch

from . import _lib
from . import device as D

_PLANS = collections.OrderedDict()
MAX_PLANS = 16          # geometry plans kept (least recently used dropped first)


def _rolled_dm(log_cents, scatter):
    """The 2N-1 distinct weights of utils._get_rolled_weights (utils.py:464-488) for a
    Normal(0, scatter) CDF: row j of the matrix holds dm[k - j + N - 1] for bins k."""
    from scipy.special import ndtr
    num = log_cents.size
    dx = np.diff(log_cents)
    if not np.allclose(dx, dx[0]):
        raise ValueError("`get_scatter_weights` only works if `uniform_cents` are uniformly spaced!")
    dx = dx[0]
    dx = dx / 2.0 + np.arange(num) * dx
    dx = np.concatenate([-dx[::-1], dx])
    return np.diff(ndtr(dx / scatter))                          # (2N-1,)


def _rolled_weights(log_cents, scatter):
    """utils._get_rolled_weights (utils.py:464-488) as a host matrix."""
    dm = _rolled_dm(log_cents, scatter)
    num = log_cents.size
    j = np.arange(num)[:, None]
    k = np.arange(num)[None, :]
    return dm[k - j + num - 1]


def _transforms(pts, simplices):
    """The affine barycentric transforms of scipy.spatial.Delaunay.transform (S, 3, 2): rows 0-1
    the inverse of [v0 - v2, v1 - v2], row 2 the vertex v2 -- in closed form for 2x2 (the same
    values to ~1e-14 relative; scipy solves each by LAPACK); NaN for a degenerate simplex, as
    scipy marks one whose reciprocal condition number is below machine epsilon."""
    P = pts[simplices]
    a = P[:, 0, 0] - P[:, 2, 0]
    b = P[:, 1, 0] - P[:, 2, 0]
    c = P[:, 0, 1] - P[:, 2, 1]
    d = P[:, 1, 1] - P[:, 2, 1]
    det = a * d - b * c
    T = np.empty((simplices.shape[0], 3, 2))
    with np.errstate(divide='ignore', invalid='ignore'):
        T[:, 0, 0] = d / det
        T[:, 0, 1] = -b / det
        T[:, 1, 0] = -c / det
        T[:, 1, 1] = a / det
    T[:, 2, :] = P[:, 2, :]
    n1 = np.maximum(np.abs(a) + np.abs(c), np.abs(b) + np.abs(d))   # 1-norm of the matrix
    with np.errstate(divide='ignore', invalid='ignore'):        # degenerate triangles: det = 0
        ni = np.maximum(np.abs(d) + np.abs(c), np.abs(b) + np.abs(a)) / np.abs(det)
        bad = ~(1.0 / (n1 * ni) >= np.finfo(float).eps)
    T[bad] = np.nan
    return T


def _simplex_g(tri, T=None):
    """Affine-invariant cross-boundary directions of each triangle (Clough-Tocher; see
    P. Alfeld, CAGD 1 (1984) 169; G. Farin, CAGD 3 (1986) 83): for edge k (opposite
    vertex k) the direction is V4' - V4, the neighbour's centroid minus this
    centroid, expressed through the barycentric coordinates c of V4'."""
    S = tri.simplices
    P = tri.points
    nb = tri.neighbors
    T = tri.transform if T is None else T
    g = np.full((S.shape[0], 3), -0.5)
    cent = P[S].mean(axis=1)
    for k in range(3):
        it = nb[:, k]
        ok = it >= 0
        y = cent[it[ok]]
        Tm = T[ok, :2, :]
        r = T[ok, 2, :]
        c01 = np.einsum('nij,nj->ni', Tm, y - r)
        c = np.concatenate([c01, 1.0 - c01.sum(axis=1, keepdims=True)], axis=1)
        if k == 0:
            gk = (2 * c[:, 2] + c[:, 1] - 1) / (2 - 3 * c[:, 2] - 3 * c[:, 1])
        elif k == 1:
            gk = (2 * c[:, 0] + c[:, 2] - 1) / (2 - 3 * c[:, 0] - 3 * c[:, 2])
        else:
            gk = (2 * c[:, 1] + c[:, 0] - 1) / (2 - 3 * c[:, 1] - 3 * c[:, 0])
        g[ok, k] = gk
    return g


def build_geometry(mtot, mrat, refine=4, kd_workers=-1):
    """Host geometry of the scatter step for one (mtot, mrat) grid (numpy/scipy and the library's
    host routines only, so it can be built in a worker process): the Delaunay triangulation of the
    (log10 m1, log10 m2) nodes, its level schedule, cross-boundary directions and affine
    transforms, and the bilinear back-interpolation.  The regular grid's point location and
    nearest nodes are made on the device from these (:meth:`ScatterPlan.device`).  Returns a
    dictionary of arrays and sizes for :class:`ScatterPlan` (``kd_workers`` is accepted for
    callers of the earlier host location and ignored)."""
    import ctypes as C
    from scipy.spatial import Delaunay
    mtot = np.asarray(mtot, dtype=float)
    mrat = np.asarray(mrat, dtype=float)
    M, Q = mtot.size, mrat.size
    m1 = mtot[:, None] / (1.0 + mrat[None, :])
    m2 = mtot[:, None] - m1
    pts = np.stack([np.log10(m1.ravel()), np.log10(m2.ravel())], axis=-1)
    G = M * refine
    lo = 0.9 * mtot[0] * mrat[0] / (1.0 + mrat[0])
    hi = mtot[-1] * (1.0 + mrat[0]) / mrat[0]
    mextr = np.array([min(lo, hi), max(lo, hi)])
    mgrid = np.log10(np.logspace(*np.log10(mextr), G))
    tri = Delaunay(pts)
    indptr, indices = tri.vertex_neighbor_vertices
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    N = pts.shape[0]
    # level schedule: a vertex may update once all lower-indexed neighbours have
    level = np.zeros(N, dtype=np.int32)
    _lib.call("hd_ct_levels", indptr.ctypes.data_as(C.c_void_p), indices.ctypes.data_as(C.c_void_p), N,
              level.ctypes.data_as(C.c_void_p))
    nlev = int(level.max()) + 1
    # the sweep runs eight lanes per vertex of a level; entries past a vertex's eighth go to an
    # overflow area
    widths = np.bincount(level, minlength=nlev)
    max_level_width = int(widths.max())
    n_ovf = int(np.maximum(np.diff(indptr) - 8, 0).sum())
    # pipelined sweeps: sweep s + 1 may run sweep_period levels behind sweep s (one more than the
    # largest level difference across an edge); a step then holds the levels l, l + P, ...
    src = np.repeat(np.arange(N), np.diff(indptr))
    dl = level[indices] - level[src]
    sweep_period = int(dl.max()) + 1 if dl.size else 1
    max_step_width = int(max(widths[r::sweep_period].sum() for r in range(sweep_period)))
    lorder = np.argsort(level, kind='stable')
    lptr = np.concatenate([[0], np.cumsum(np.bincount(level, minlength=nlev))])
    # bilinear back-interpolation (RegularGridInterpolator 'linear') at the nodes
    ii = np.clip(np.searchsorted(mgrid, pts[:, 0]) - 1, 0, G - 2)
    jj = np.clip(np.searchsorted(mgrid, pts[:, 1]) - 1, 0, G - 2)
    wx = (pts[:, 0] - mgrid[ii]) / (mgrid[ii + 1] - mgrid[ii])
    wy = (pts[:, 1] - mgrid[jj]) / (mgrid[jj + 1] - mgrid[jj])
    back_idx = np.stack([ii * G + jj, ii * G + jj + 1, (ii + 1) * G + jj, (ii + 1) * G + jj + 1], axis=-1)
    back_w = np.stack([(1 - wx) * (1 - wy), (1 - wx) * wy, wx * (1 - wy), wx * wy], axis=-1)
    T = _transforms(pts, tri.simplices)
    host = dict(
        pts=pts, nbr_ptr=indptr, nbr_idx=indices,
        levels_ptr=lptr.astype(np.int32), levels_idx=lorder.astype(np.int32),
        simplices=tri.simplices.astype(np.int32), simplex_g=_simplex_g(tri, T), transform=T, mgrid=mgrid,
        back_idx=back_idx.astype(np.int32), back_w=back_w)
    return dict(host=host, M=M, Q=Q, G=G, N=N, n_levels=nlev, max_level_width=max_level_width, n_ovf=n_ovf,
                sweep_period=sweep_period, max_step_width=max_step_width, mgrid=mgrid)


class ScatterPlan:
    """Geometry of the scatter step for one (mtot, mrat) grid; the Gaussian-CDF convolution matrix
    for a given scatter is made on demand (:meth:`P_for`).  ``geom``: the output of
    :func:`build_geometry` if it was built elsewhere."""

    def __init__(self, mtot, mrat, refine=4, geom=None):
        g = build_geometry(mtot, mrat, refine) if geom is None else geom
        self.M, self.Q, self.G, self.N = g['M'], g['Q'], g['G'], g['N']
        self.n_levels = g['n_levels']
        self.max_level_width = g['max_level_width']
        self.n_ovf = g['n_ovf']
        self.sweep_period = g['sweep_period']
        self.max_step_width = g['max_step_width']
        self.host = g['host']
        self.mgrid = g['mgrid']
        self._dev = None
        self._ready = None
        self._P = {}

    def device(self):
        """Device copies of the geometry for work on the current stream.  A stream other than the
        uploading one waits for the uploads (an event); each use marks the tensors as used by the
        current stream, so that a plan dropped from the cache is not recycled while another
        stream still reads it."""
        cur = torch.cuda.current_stream()
        if self._dev is None:
            h = self.host
            d = {}
            for k, v in h.items():
                d[k] = D.to_dev(v, dtype=torch.int32 if v.dtype == np.int32 else torch.float64)
            # the regular grid's simplices, barycentric coordinates and nearest nodes
            G2 = self.G * self.G
            d['q_simplex'] = D.empty(G2, dtype=torch.int32)
            d['q_bary'] = D.empty((G2, 3))
            d['q_nearest'] = D.empty(G2, dtype=torch.int32)
            _lib.call("hd_ct_locate", D.ptr(d['pts']), self.N, D.ptr(d['simplices']), D.ptr(d['transform']),
                      int(h['simplices'].shape[0]), D.ptr(d['mgrid']), self.G, D.ptr(d['q_simplex']),
                      D.ptr(d['q_bary']), D.ptr(d['q_nearest']), D.stream_ptr())
            self._dev = d
            self._ready = torch.cuda.Event()
            self._ready.record(cur)
        else:
            cur.wait_event(self._ready)
        for t in self._dev.values():
            t.record_stream(cur)
        return self._dev

    def P_for(self, scatter, keep=8):
        """P = (W W)^T = W^T W^T for the Gaussian-CDF weights of `scatter` [dex], made on the device
        on the current stream (the last `keep` scatter values are cached; a later user on another
        stream waits for the product)."""
        key = float(scatter)
        cur = torch.cuda.current_stream()
        ent = self._P.get(key)
        if ent is None:
            G = self.G
            dm = D.to_dev(_rolled_dm(self.mgrid, scatter))
            wT = D.empty((G, G))
            _lib.call("hd_scatter_weights_T", D.ptr(dm), G, D.ptr(wT), D.stream_ptr())
            P = D.empty((G, G))
            _lib.call("hd_gemm_f64", D.ptr(wT), D.ptr(wT), D.ptr(P), G, G, G, D.stream_ptr())
            ev = torch.cuda.Event()
            ev.record(cur)
            if len(self._P) >= keep:
                self._P.pop(next(iter(self._P)))
            self._P[key] = (P, ev)
            return P
        P, ev = ent
        cur.wait_event(ev)
        P.record_stream(cur)
        return P


_POOL = None


def _pool_init():
    # one thread per worker process: the pool supplies the parallelism
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def geometry_pool(workers):
    """A process pool (spawned interpreters, reused across calls) for :func:`build_geometry`."""
    global _POOL
    import concurrent.futures
    import multiprocessing
    if _POOL is None or _POOL._max_workers != workers:
        if _POOL is not None:
            _POOL.shutdown(wait=True, cancel_futures=True)
        _POOL = concurrent.futures.ProcessPoolExecutor(max_workers=workers,
                                                       mp_context=multiprocessing.get_context("spawn"),
                                                       initializer=_pool_init)
        atexit.register(_POOL.shutdown, wait=True, cancel_futures=True)
    return _POOL


def plan_key(mtot, mrat):
    return (np.asarray(mtot, dtype=float).tobytes(), np.asarray(mrat, dtype=float).tobytes())


def has_plan(mtot, mrat):
    return plan_key(mtot, mrat) in _PLANS


def put_plan(plan, mtot, mrat):
    """Insert a plan built elsewhere (e.g. ahead of time in a worker thread) into the cache."""
    key = plan_key(mtot, mrat)
    while len(_PLANS) >= MAX_PLANS:
        _PLANS.popitem(last=False)
    _PLANS[key] = plan


def get_plan(mtot, mrat):
    """Cached geometry plan of one (mtot, mrat) grid."""
    key = plan_key(mtot, mrat)
    plan = _PLANS.get(key)
    if plan is None:
        plan = ScatterPlan(mtot, mrat)
        while len(_PLANS) >= MAX_PLANS:
            _PLANS.popitem(last=False)
        _PLANS[key] = plan
    else:
        _PLANS.move_to_end(key)
    return plan


#: run the Gauss-Seidel sweeps pipelined (sweep s + 1 behind sweep s by the plan's period) when they
#: fit the device; False runs them one after another (the same results bit for bit)
PIPELINED_SWEEPS = True


def add_scatter_to_masses_device(mtot, mrat, dens, scatter, tol=1e-6, maxiter=400, return_grads=False):
    """In place on the (M, Q, Z) device density; returns the per-slice Gauss-Seidel sweep counts (0 = no
    convergence within maxiter, as the reference's estimator reports).  ``return_grads``: also the
    Clough-Tocher node gradients of every slice, (Z, N, 2) on the device (scipy's
    ``CloughTocher2DInterpolator(...).grad`` of that slice, nodes in C order of (M, Q))."""
    plan = get_plan(mtot, mrat)
    d = dict(plan.device())
    d['P'] = plan.P_for(scatter)
    M, Q, Z = dens.shape
    assert (M, Q) == (plan.M, plan.Q)
    period = plan.sweep_period if PIPELINED_SWEEPS else 0
    ws = D.workspace(_lib.size("hd_scatter_workspace_size", M, Q, Z, plan.G, plan.n_ovf, plan.n_levels, period))
    iters = D.empty(Z + 1, dtype=torch.int32)
    from .engine import _stage
    with _stage("scatter"):
        _scatter_call(dens, M, Q, Z, plan, d, tol, maxiter, ws, iters, period)
    if not return_grads:
        return iters
    # the gradients are the workspace's first region, at its first 256-byte boundary
    off = (-ws.data_ptr()) % 256
    grads = ws[off:off + Z * plan.N * 16].view(torch.float64).view(Z, plan.N, 2)
    return iters, grads


def _scatter_call(dens, M, Q, Z, plan, d, tol, maxiter, ws, iters, period):
    _lib.call("hd_scatter", D.ptr(dens), M, Q, Z, plan.G, D.ptr(d['pts']), D.ptr(d['nbr_ptr']), D.ptr(d['nbr_idx']),
              D.ptr(d['levels_ptr']), D.ptr(d['levels_idx']), plan.n_levels, plan.max_level_width, plan.n_ovf,
              int(period), plan.max_step_width,
              D.ptr(d['simplices']),
              D.ptr(d['simplex_g']), D.ptr(d['q_simplex']), D.ptr(d['q_bary']), D.ptr(d['q_nearest']),
              D.ptr(d['P']), D.ptr(d['back_idx']), D.ptr(d['back_w']), float(tol), int(maxiter), D.ptr(ws),
              ws.numel(), D.ptr(iters), D.stream_ptr())
