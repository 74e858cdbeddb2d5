"""GPU parity: every kernel of the path against the FP64 CPU oracle on the same inputs.

Tolerances: FP64 stage-A kernels 1e-9 relative (different libm pow/log ulps and
one table-walk tie rule); realization with injected counts:
h_c within 1e-5 relative (observed ~1e-13), loud-source
bin indices bit-exact.
"""
import numpy as np
import pytest

from holodeck_amd import configs

pytestmark = pytest.mark.gpu

SMALL = (17, 15, 13)


def _close(a, b, rtol, atol=0.0, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin), f"{what}: finite pattern differs"
    err = np.abs(a[fin] - b[fin])
    tol = atol + rtol * np.abs(b[fin])
    bad = err > tol
    assert not bad.any(), f"{what}: {bad.sum()} / {bad.size} outside rtol={rtol}; worst {np.max(err / (np.abs(b[fin]) + 1e-300))}"


def _stageA(spec, gpu, oracle_lib):
    from oracle.cfg import oracle_cfg
    from holodeck_amd import device as D
    sam = configs.make_sam(spec, scatter=0.0)
    hard = configs.make_hard(spec, sam)
    cfg = oracle_cfg(spec)
    cfg['mmbulge']['scatter_dex'] = 0.0
    o = oracle_lib.number_h2fdf(cfg)
    fe = configs.fobs_gw_edges(spec)
    number, h2fdf, rf = sam._number_h2fdf_device(fe, hard)
    return sam, hard, o, D.to_host(number), D.to_host(h2fdf), D.to_host(rf)


@pytest.mark.parametrize("which", ["config1", "config2", "config2_many_f", "config2_more_f"])
def test_stageA_parity(which, gpu, oracle_lib):
    """config2_many_f: 101 redshift edges x 100 frequencies (long bracket walks, many targets
    per step); config2_more_f: 200 frequencies, whose recorded steps (Z x F) no longer fit the
    evolution kernel's LDS, so they are kept in global memory."""
    print(f"[stageA {which}]", flush=True)
    if which == "config1":
        spec = configs.spec_config1(shape=SMALL, nfreq=7)
    elif which == "config2":
        spec = configs.spec_config2(shape=SMALL, nfreq=7)
    elif which == "config2_many_f":
        spec = configs.spec_config2(shape=(7, 6, 101), nfreq=100)
    else:
        spec = configs.spec_config2(shape=(7, 6, 101), nfreq=200)
    sam, hard, o, number, h2fdf, rf = _stageA(spec, gpu, oracle_lib)
    _close(sam.static_binary_density, o['dens'], 1e-12, what="density")
    if 'norm_log10' in o:
        from holodeck_amd import device as D
        _close(D.to_host(hard._norm_log10_dev), o['norm_log10'], 1e-12, what="norm_log10")
    _close(rf, o['redz_final'], 1e-9, what="redz_final")
    F = number.shape[-1]
    _close(number, o['number'].reshape(-1, F), 1e-9, atol=0.0, what="number")
    _close(h2fdf, o['h2fdf'].reshape(-1, F), 1e-9, what="h2fdf")


@pytest.mark.parametrize("i0", [0, 44, 88])
def test_stageA_full_size_mass_rows(i0, gpu, oracle_lib):
    """The full-size grid in one run (91 x 81 x 101 edges, 40 frequencies, 2PL evolution; no mass
    scatter, whose kernels have tests of their own) against the oracle on three consecutive
    total-mass rows, with every ratio, redshift and frequency: density, Brent norms, redshift at
    the targets per corner, number and strain in the two mass bins between the rows.  The
    oracle's evolution is independent per (M, Q), so the rows are run alone."""
    print(f"[stageA 91x81x101 F=40 mass rows {i0}..{i0 + 2}]", flush=True)
    from oracle import pipeline as P
    from oracle.cfg import oracle_cfg
    from holodeck_amd import device as D
    spec = configs.spec_config2(nfreq=40)
    sam = configs.make_sam(spec, scatter=0.0)
    hard = configs.make_hard(spec, sam)
    number, h2fdf, rf = sam._number_h2fdf_device(configs.fobs_gw_edges(spec), hard)
    cfg = oracle_cfg(spec)
    cfg['mmbulge']['scatter_dex'] = 0.0
    c = cfg['cosmo']
    cosmo = P.OracleCosmo(h=c['h'], Om0=c['Om0'], size=c['size'], zmax=c['zmax'])
    mtot, mrat, redz = P.edges_from_cfg(cfg)
    M, Q, Z = len(mtot), len(mrat), len(redz)
    assert (M, Q, Z) == (91, 81, 101)
    dens, zprime, gmt = P.static_density(mtot, mrat, redz, cfg, cosmo)
    dens = np.array(dens)
    dens[zprime < 0.0] = 0.0
    _close(sam.static_binary_density, dens, 1e-12, what="density")
    rows = slice(i0, i0 + 3)
    hd = cfg['hard']
    mt2, mr2 = np.meshgrid(mtot[rows], mrat, indexing='ij')
    nl = P.find_2pwl_norm(hd['time'], mt2.ravel(), mr2.ravel(), hd['sepa_init'], hd['rchar'], hd['gamma_inner'],
                          hd['gamma_outer'], hd['num_steps']).reshape(mt2.shape)
    _close(D.to_host(hard._norm_log10_dev).reshape(M, Q)[rows], nl, 1e-12, what="norm_log10")
    fge = np.asarray(cfg['fobs_gw_edges'])
    foe, foc = fge / 2.0, P.midpoints(fge) / 2.0
    rf_o, dn_o = P.dbn_2pwl(foc, hd['sepa_init'], hd['num_steps'], 10.0 ** nl, hd['rchar'], hd['gamma_inner'],
                            hd['gamma_outer'], dens[rows], mtot[rows], mrat, redz, gmt[rows], cosmo)
    F = len(foc)
    _close(D.to_host(rf).reshape(M, Q, Z, F)[rows], rf_o, 1e-9, what="redz_final")
    edges = [mtot[rows], mrat, redz, foe]
    n_o = P.integrate_3dx1d(edges, dn_o)
    h_o = P.h2fdf_redz(edges, rf_o, cosmo)
    bins = slice(i0, i0 + 2)
    _close(D.to_host(number).reshape(M - 1, Q - 1, Z - 1, F)[bins], n_o, 1e-9, atol=1e-12 * np.abs(n_o).max(),
           what="number")
    _close(D.to_host(h2fdf).reshape(M - 1, Q - 1, Z - 1, F)[bins], h_o, 1e-9, what="h2fdf")


def _rand_counts(number, R, seed):
    rng = np.random.default_rng(seed)
    lam = np.minimum(number, 1e6)
    return rng.poisson(lam[..., None] * np.ones(R)).astype(float)


def test_realize_counts_sorted(gpu, oracle_lib):
    print("[realize counts sorted]", flush=True)
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    rng = np.random.default_rng(1)
    B, F, R, L = 3000, 5, 7, 4
    h2 = 10.0 ** rng.uniform(-35, -25, size=(B, F))
    h2[rng.random((B, F)) < 0.05] = 0.0
    number = 10.0 ** rng.uniform(-4, 1, size=(B, F))
    counts = _rand_counts(number, R, 2)
    order = oracle_lib.order_desc(h2[:, 0])
    for skip in (False, True):
        ss_o, idx_o, bg_o, _ = oracle_lib.loudest_counts(counts, h2, order, L, skip_zero_h=skip)
        ss, idx, bg, _ = engine.realize(D.to_dev(number), D.to_dev(h2), R, L, _lib.HD_SEL_SORTED,
                                        order=D.to_dev(order), counts=D.to_dev(counts),
                                        flags=_lib.HD_FLAG_SKIP_ZERO_H if skip else 0)
        assert np.array_equal(D.to_host(idx), idx_o)
        _close(D.to_host(ss), ss_o, 0.0, what="hc2ss")
        _close(D.to_host(bg), bg_o, 1e-12, what="hc2bg")


def test_realize_counts_argmax_and_none(gpu, oracle_lib):
    print("[realize counts argmax/none]", flush=True)
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    rng = np.random.default_rng(3)
    B, F, R = 2500, 4, 6
    h2 = 10.0 ** rng.uniform(-35, -25, size=(B, F))
    number = 10.0 ** rng.uniform(-4, 0.5, size=(B, F))
    counts = _rand_counts(number, R, 4)
    par = rng.uniform(1, 2, size=(3, B))
    ss_o, bg_o, idx_o, bgp_o = oracle_lib.ss_bg_counts(counts, h2, par=par)
    pdev = D.to_dev(np.repeat(par[:, :, None], F, axis=2))
    ss, idx, bg, bgp = engine.realize(D.to_dev(number), D.to_dev(h2), R, 1, _lib.HD_SEL_ARGMAX,
                                      counts=D.to_dev(counts), par=pdev)
    assert np.array_equal(D.to_host(idx)[..., 0], idx_o)
    _close(D.to_host(ss)[..., 0], ss_o, 0.0, what="ss argmax")
    _close(D.to_host(bg), bg_o, 1e-10, what="bg argmax")
    _close(D.to_host(bgp), bgp_o, 1e-10, what="bgpar argmax")
    gwb_o = oracle_lib.poisson_gwb_counts(counts, h2)
    _, _, bgn, _ = engine.realize(D.to_dev(number), D.to_dev(h2), R, 0, _lib.HD_SEL_NONE, counts=D.to_dev(counts))
    _close(D.to_host(bgn), gwb_o, 1e-12, what="poisson gwb")


def test_order_matches_stable_argsort(gpu, oracle_lib):
    print("[order]", flush=True)
    from holodeck_amd import engine
    from holodeck_amd import device as D
    rng = np.random.default_rng(5)
    for B in (1, 7, 4096, 4097, 100003):
        key = 10.0 ** rng.uniform(-40, -20, size=(B, 3))
        key[rng.random((B, 3)) < 0.2] = 0.0          # many ties
        key[rng.random((B, 3)) < 0.05] = 1e-30        # more ties
        o = engine.order_desc(D.to_dev(key), col=1)
        assert np.array_equal(D.to_host(o), oracle_lib.order_desc(key[:, 1])), B


def test_pipeline_counts_end_to_end(gpu, oracle_lib):
    """gwb with injected counts vs the oracle pipeline: h_c 1e-5 relative (observed ~1e-12), indices exact."""
    print("[pipeline injected counts]", flush=True)
    spec = configs.spec_config2(shape=SMALL, nfreq=6)
    sam, hard, o, number, h2fdf, rf = _stageA(spec, gpu, oracle_lib)
    F = number.shape[-1]
    R, L = 5, 3
    counts = _rand_counts(o['number'].reshape(-1, F), R, 9)
    order = oracle_lib.order_desc(o['h2fdf'].reshape(-1, F)[:, 0])
    ss_o, idx_o, bg_o, _ = oracle_lib.loudest_counts(counts, o['h2fdf'].reshape(-1, F), order, L)
    hc_ss, hc_bg = sam.gwb(configs.fobs_gw_edges(spec), hard=hard, realize=R, loudest=L, counts=counts)
    _close(hc_bg, np.sqrt(bg_o), 1e-5, what="hc_bg")
    _close(hc_ss, np.sqrt(ss_o), 1e-5, what="hc_ss")


def test_scatter_and_density_full_config1(gpu, oracle_lib):
    """Density with M-Mbulge scatter at the full 61x51x41 grid vs the oracle's scipy-based restatement."""
    print("[scatter full config1]", flush=True)
    from oracle.cfg import oracle_cfg
    from oracle.scatter import add_scatter_to_masses
    from holodeck_amd import device as D
    spec = configs.spec_config1()
    sam = configs.make_sam(spec)
    dens = sam.static_binary_density
    iters = D.to_host(sam._scatter_iters)
    assert np.all(iters[:-1] > 0) and iters[-1] == 0, iters
    cfg = oracle_cfg(spec)
    c = cfg['cosmo']
    cos = oracle_lib.OracleCosmo(h=c['h'], Om0=c['Om0'])
    mtot, mrat, redz = oracle_lib.edges_from_cfg(cfg)
    d0, zp, _ = oracle_lib.static_density(mtot, mrat, redz, cfg, cos)
    do = add_scatter_to_masses(mtot, mrat, d0, cfg['mmbulge']['scatter_dex'])
    do[zp < 0] = 0.0
    scale = np.max(np.abs(do))
    _close(dens, do, 1e-8, atol=1e-13 * scale, what="scattered density")


@pytest.mark.parametrize("pipelined", [True, False])
def test_ct_gradients_bit_exact_vs_scipy(gpu, pipelined):
    """The Gauss-Seidel gradient estimate of every slice equals scipy's CloughTocher2DInterpolator
    gradients (tol 1e-6, maxiter 400) bit for bit, at the full 91 x 81 mass grid: five slices of the
    static density, one with a zeroed quadrant (zero numerators, signed zeros), and the density scaled
    to 1e-295 and 1e280 (the divisions' exact fallback outside the fast step's range).  Both with the
    sweeps pipelined (sweep s + 1 a period of levels behind sweep s) and one after another; and with
    maxiter = 3, where most slices stop unconverged (scipy's gradients after three sweeps, count 0)."""
    import scipy.interpolate
    import torch
    from holodeck_amd import device as D, scatter
    spec = configs.spec_config2()
    sam = configs.make_sam(spec, scatter=0.0)
    dens0, _, _ = sam._density_device()
    base = D.to_host(dens0)
    Zall = base.shape[-1]
    sl = [base[:, :, z] for z in (0, Zall // 4, Zall // 2, 3 * Zall // 4, Zall - 1)]
    zq = base[:, :, Zall // 2].copy()
    zq[: zq.shape[0] // 2, : zq.shape[1] // 2] = 0.0
    sl += [zq, base[:, :, Zall // 2] * 1e-295, base[:, :, Zall // 2] * 1e280]
    dens = np.stack(sl, axis=-1)
    m1 = sam.mtot[:, None] / (1.0 + sam.mrat[None, :])
    m2 = sam.mtot[:, None] - m1
    pts = (np.log10(m1.ravel()), np.log10(m2.ravel()))
    keep = scatter.PIPELINED_SWEEPS
    scatter.PIPELINED_SWEEPS = pipelined
    try:
        for maxiter in (400, 3):
            print(f"[ct gradients {dens.shape} pipelined={pipelined} maxiter={maxiter}]", flush=True)
            d = D.to_dev(dens)
            iters, grads = scatter.add_scatter_to_masses_device(sam.mtot, sam.mrat, d, 0.28, maxiter=maxiter,
                                                                return_grads=True)
            torch.cuda.synchronize()
            g = D.to_host(grads)
            it = D.to_host(iters)
            for z in range(dens.shape[-1]):
                want = np.asarray(scipy.interpolate.CloughTocher2DInterpolator(pts, dens[:, :, z].ravel(),
                                                                               maxiter=maxiter).grad)
                want = want.reshape(-1, 2)
                diff = np.flatnonzero(g[z].view(np.int64) != want.view(np.int64))
                assert diff.size == 0, (maxiter, z, it[z], diff.size, diff[:4], g[z].ravel()[diff[:4]],
                                        want.ravel()[diff[:4]])
                assert (it[z] > 0) if maxiter == 400 else (0 <= it[z] <= 3), (maxiter, z, it[z])
            if maxiter == 3:   # the stop at maxiter is taken (slices that need more sweeps report 0)
                assert np.any(it[:-1] == 0), it
    finally:
        scatter.PIPELINED_SWEEPS = keep


@pytest.mark.parametrize("shape", [(17, 15), (61, 51), (91, 81)])
def test_ct_locate_matches_scipy(gpu, shape):
    """hd_ct_locate on the scatter's regular grid against scipy: every point scipy's
    find_simplex places is placed in a simplex that holds it (the same one, or on a shared edge a
    neighbour: both barycentric tests pass), with barycentric coordinates to 1e-12; points outside
    the hull get -1 as in scipy; the nearest node is cKDTree's (or one at the same distance)."""
    from scipy.spatial import Delaunay, cKDTree
    from holodeck_amd import scatter
    from holodeck_amd import device as D
    from holodeck_amd.constants import MSOL
    M, Q = shape
    print(f"[ct locate {shape}]", flush=True)
    mtot = np.logspace(np.log10(1e4 * MSOL), np.log10(1e12 * MSOL), M)
    mrat = np.logspace(-3, 0, Q)
    plan = scatter.ScatterPlan(mtot, mrat)
    d = plan.device()
    simp, bary, near = D.to_host(d['q_simplex']), D.to_host(d['q_bary']), D.to_host(d['q_nearest'])
    pts, mgrid, G = plan.host['pts'], plan.mgrid, plan.G
    X, Y = np.meshgrid(mgrid, mgrid, indexing='ij')
    xi = np.stack([X.ravel(), Y.ravel()], axis=-1)
    tri = Delaunay(pts)
    want = tri.find_simplex(xi)
    assert np.array_equal(simp < 0, want < 0), (np.sum(simp < 0), np.sum(want < 0))
    ins = want >= 0
    diff = ins & (simp != want)
    T = tri.transform
    for k in np.flatnonzero(diff):   # a different simplex only where both hold the point
        c = T[simp[k], :2] @ (xi[k] - T[simp[k], 2])
        c = np.append(c, 1 - c.sum())
        assert np.all(c >= -1e-13) and np.all(c <= 1 + 1e-13), (k, c)
    same = ins & ~diff
    c01 = np.einsum('nij,nj->ni', T[want[same], :2, :], xi[same] - T[want[same], 2, :])
    assert np.allclose(bary[same, :2], c01, rtol=0, atol=1e-12)
    dist, nn = cKDTree(pts).query(xi)
    dgot = np.hypot(*(xi - pts[near]).T)
    assert np.all((near == nn) | (np.abs(dgot - dist) <= 1e-15 * np.maximum(dist, 1e-300)))


def test_stageA_full_config1_with_scatter(gpu, oracle_lib):
    """number and h2fdf at the full config-1 workload (61x51x41, 20 frequencies, GW-only)."""
    print("[stageA full config1]", flush=True)
    from oracle.cfg import oracle_cfg
    from holodeck_amd import device as D
    spec = configs.spec_config1()
    sam = configs.make_sam(spec)
    hard = configs.make_hard(spec, sam)
    o = oracle_lib.number_h2fdf(oracle_cfg(spec))
    number, h2fdf, rf = sam._number_h2fdf_device(configs.fobs_gw_edges(spec), hard)
    F = number.shape[-1]
    n_o = o['number'].reshape(-1, F)
    _close(D.to_host(rf), o['redz_final'], 1e-9, what="redz_final")
    _close(D.to_host(number), n_o, 1e-7, atol=1e-12 * n_o.max(), what="number")
    h_o = o['h2fdf'].reshape(-1, F)
    _close(D.to_host(h2fdf), h_o, 1e-9, what="h2fdf")


def test_mass_stellar(gpu, oracle_lib):
    """mass_stellar (sam.py:250-278) on the device against the oracle's M-Mbulge relation."""
    print("[mass_stellar config2 small]", flush=True)
    from oracle.cfg import oracle_cfg
    from oracle.pipeline import mstar_from_mbh
    spec = configs.spec_config2(shape=SMALL, nfreq=7)
    sam = configs.make_sam(spec, scatter=0.0)
    mp, mr, mt, rz = sam.mass_stellar()
    mmb = oracle_cfg(spec)['mmbulge']
    m1 = sam.mtot[:, None] / (1.0 + sam.mrat[None, :])
    m2 = sam.mtot[:, None] - m1
    s1 = np.broadcast_to(mstar_from_mbh(m1, mmb)[..., None], mp.shape)
    s2 = np.broadcast_to(mstar_from_mbh(m2, mmb)[..., None], mp.shape)
    _close(mp, s1, 1e-12, what="mstar_pri")
    _close(mr, s2 / s1, 1e-12, what="mstar_rat")
    _close(mt, s1 + s2, 1e-12, what="mstar_tot")
    assert rz.shape == mp.shape and np.array_equal(rz[0, 0], sam.redz)
    # secondary never heavier than primary (the relation is monotonic and q <= 1)
    assert np.all(mr <= 1.0 + 1e-15)
