"""Pins of the CPU oracle (runs without a GPU).

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails it:
* GW / Kepler formulas and constants: the reference's own known-answer
  vectors (holodeck/tests/test_utils.py:60-72), as tests/golden/gw_formulas.json;
* Brent iteration: scipy.optimize.brentq on the same lifetime function (same
  published algorithm as the reference's cython_optimize.brentq);
* 2PL lifetime: GW-only closed form and second-order convergence of the
  300-step trapezoid;
* cosmology: mpmath quadrature of the Friedmann integrals at 30 digits;
* 3D x 1D integration: exactness on functions linear in log10 M, q, z;
* realization walks: an independent instance-expansion implementation;
* GW-only dynamic binary number: a vectorised numpy restatement of the
  reference's own Python version (sam.py:618-676);
* 2PL dynamic binary number: an independent continuous-time integration
  (scipy.integrate.quad) of the same hardening law and exact cosmology.
"""
import json
import os

import numpy as np
import pytest

import oracle
from oracle.pipeline import OracleCosmo, MSOL, PC, GYR, YR, SPLC, MPC

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gw_formulas.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as fh:
        return {k: np.array(v) for k, v in json.load(fh).items()}


def _peters_fe(e):
    return (1 + (73 / 24) * e ** 2 + (37 / 96) * e ** 4) / (1 - e ** 2) ** 3.5


def test_chirp_mass_and_strain_golden(gold):
    m1, m2 = gold['m1'], gold['m2']
    mt = m1 + m2
    q = np.minimum(m1, m2) / np.maximum(m1, m2)
    np.testing.assert_allclose(oracle.chirp_mass_mtmr(mt, q), gold['mc'], rtol=2e-8, atol=0)
    hs = oracle.gw_strain_source(gold['mc'], gold['dc'], 1.0 / YR)
    np.testing.assert_allclose(hs, gold['hs'], rtol=2e-8, atol=0)


def test_hard_gw_and_residence_time_golden(gold):
    L = oracle.lib()
    m1, m2, aa, ee = gold['m1'], gold['m2'], gold['aa'], gold['ee']
    mt = m1 + m2
    q = np.minimum(m1, m2) / np.maximum(m1, m2)
    dadt = np.array([L.or_hard_gw(a, b, c) for a, b, c in zip(mt, q, aa)]) * _peters_fe(ee)
    # the Cython module's G (6.6742999e-8) differs from astropy's by 1.5e-8; G^3 -> 4.5e-8
    np.testing.assert_allclose(dadt, gold['dadt'], rtol=1e-7, atol=0)
    # circular GW residence time f/(df/dt) = (5/96)(G Mc/c^3)^(-5/3)(2 pi f)^(-8/3)
    tres = np.array([L.or_gw_tres(a, b, 1.0 / YR) for a, b in zip(mt, q)])
    np.testing.assert_allclose(tres, gold['tau'], rtol=1e-7, atol=0)
    # df/dt = -(3/2)(f/a) da/dt with a = Kepler separation at f (eccentric golden values)
    f = 1.0 / YR
    a = np.array([L.or_kepler_sepa(m, f) for m in mt])
    dadt_e = np.array([L.or_hard_gw(x, y, z) for x, y, z in zip(mt, q, a)]) * _peters_fe(ee)
    np.testing.assert_allclose(-1.5 * (f / a) * dadt_e, gold['dfdt'], rtol=1e-7, atol=0)


def _brent_case(mt, mr, time, sepa_init, rchar, gi, go, nsteps):
    from scipy.optimize import brentq
    L = oracle.lib()
    import ctypes as C
    it = C.c_int(0)
    nf = C.c_int(0)
    root = L.or_brentq_lifetime(-20.0, 20.0, 1e-3, 1e-5, 100, mt, mr, time, sepa_init, rchar, gi, go, nsteps,
                                C.addressof(it), C.addressof(nf))
    f = lambda x: L.or_lifetime_2pwl(x, mt, mr, time, sepa_init, rchar, gi, go, nsteps)  # noqa: E731
    ref, info = brentq(f, -20.0, 20.0, xtol=1e-3, rtol=1e-5, maxiter=100, full_output=True)
    return root, ref, it.value, info.iterations, nf.value, info.function_calls


@pytest.mark.parametrize("gi,go,rchar,sepa,time", [(-1.0, 2.5, 100 * PC, 1e4 * PC, 1.0 * GYR),
                                                   (-1.0, 1.5, 10 * PC, 1e3 * PC, 3.0 * GYR),
                                                   (-0.3, 2.5, 100 * PC, 1e4 * PC, 0.1 * GYR),
                                                   (-1.5, 2.5, 100 * PC, 1e4 * PC, 11.0 * GYR)])
def test_brentq_matches_scipy(gi, go, rchar, sepa, time):
    for mt in (1e4 * MSOL, 1e7 * MSOL, 1e10 * MSOL, 1e12 * MSOL):
        for mr in (1e-3, 0.1, 1.0):
            root, ref, it, it_ref, nf, nf_ref = _brent_case(mt, mr, time, sepa, rchar, gi, go, 300)
            assert root == ref, (mt, mr, root, ref)
            assert it == it_ref and nf == nf_ref


def test_lifetime_gw_limit_and_convergence():
    L = oracle.lib()
    mt, mr = 1e9 * MSOL, 0.3
    a0 = 1e-2 * PC
    # GW only (norm 1e-20 makes the 2PL term negligible): t = (a0^4 - a1^4) / (4 K)
    c = np.zeros(6)
    L.or_constants(c.ctypes.data)
    gw_const = c[0]
    K = -gw_const * mt ** 3 * mr / (1 + mr) ** 2
    a1 = 3.0 * c[4] * mt
    t_exact = (a0 ** 4 - a1 ** 4) / (4 * K)
    errs = []
    for n in (300, 600, 1200):
        t = L.or_lifetime_2pwl(-20.0, mt, mr, 0.0, a0, 10 * PC, -1.0, 2.5, n)
        errs.append(abs(t - t_exact) / t_exact)
    assert errs[0] < 2e-3
    # second-order scheme: error drops ~4x per halving of the step
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_cosmology_vs_mpmath():
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    for h, om in ((0.6933, 0.288), (0.677, 0.31)):
        c = OracleCosmo(h=h, Om0=om)
        H0 = mpmath.mpf(h) * 100 * mpmath.mpf(10) ** 5 / mpmath.mpf(MPC)
        E = lambda z: mpmath.sqrt(mpmath.mpf(om) * (1 + z) ** 3 + (1 - mpmath.mpf(om)))  # noqa: E731
        for z in (1e-3, 0.1, 1.0, 3.7, 10.0):
            dc = mpmath.mpf(SPLC) / H0 * mpmath.quad(lambda x: 1 / E(x), [0, z])
            age = mpmath.quad(lambda x: 1 / ((1 + x) * E(x)), [z, mpmath.inf]) / H0
            assert abs(float(c.comoving_distance(z)) / float(dc) - 1) < 1e-13
            assert abs(float(c.age(z)) / float(age) - 1) < 1e-13
            assert abs(float(c.dtdz(z)) * float(H0) * (1 + z) * float(E(z)) - 1) < 1e-13
            assert abs(float(c.tage_to_z(c.age(z))) - z) < 1e-10 * (1 + z)


def test_integrate_exact_on_trilinear():
    M, Q, Z, F = 6, 5, 4, 3
    mt = np.logspace(30, 40, M)
    mr = np.logspace(-3, 0, Q)
    rz = np.logspace(-3, 1, Z)
    fe = np.array([1e-9, 2e-9, 3.5e-9, 6e-9])
    lm = np.log10(mt)
    a, b, c, d = 1.3, -0.7, 2.1, 0.4
    dnum = (a + b * lm[:, None, None] + c * mr[None, :, None] + d * rz[None, None, :])[..., None] * np.ones(F)
    num = oracle.integrate_3dx1d([mt, mr, rz, fe], dnum)
    # exact integral of a trilinear function over each cell, Riemann in ln f
    lo = lambda x: x[:-1]  # noqa: E731
    hi = lambda x: x[1:]  # noqa: E731
    vol = (hi(lm) - lo(lm))[:, None, None] * (hi(mr) - lo(mr))[None, :, None] * (hi(rz) - lo(rz))[None, None, :]
    mean = a + b * 0.5 * (hi(lm) + lo(lm))[:, None, None] + c * 0.5 * (hi(mr) + lo(mr))[None, :, None] \
        + d * 0.5 * (hi(rz) + lo(rz))[None, None, :]
    expect = (mean * vol)[..., None] * np.diff(np.log(fe))
    np.testing.assert_allclose(num, expect, rtol=1e-12)


def _walk_by_expansion(counts, h2, order, L, skip_zero_h):
    """Independent implementation: expand each bin into ceil(n) unit instances along the order."""
    B, F, R = counts.shape
    ss = np.zeros((F, R, L))
    idx = -np.ones((F, R, L), dtype=np.int64)
    bg = np.zeros((F, R))
    for f in range(F):
        for r in range(R):
            n = counts[order, f, r]
            h = h2[order, f]
            keep = n >= 1
            if skip_zero_h:
                keep &= h != 0
            n, h, b = n[keep], h[keep], order[keep]
            inst = np.repeat(np.arange(n.size), np.ceil(n).astype(int))
            take = inst[:L]
            ss[f, r, :take.size] = h[take]
            idx[f, r, :take.size] = b[take]
            bg[f, r] = np.sum(n * h) - np.sum(h[take])
    return ss, idx, bg


def test_realization_walks_vs_expansion():
    rng = np.random.default_rng(11)
    B, F, R, L = 400, 3, 5, 6
    h2 = 10.0 ** rng.uniform(-35, -25, size=(B, F))
    h2[rng.random((B, F)) < 0.1] = 0.0
    counts = rng.poisson(10.0 ** rng.uniform(-2, 0.5, size=(B, F, R))).astype(float)
    order = oracle.order_desc(h2[:, 0])
    for skip in (False, True):
        ss, idx, bg, _ = oracle.loudest_counts(counts, h2, order, L, skip_zero_h=skip)
        ss2, idx2, bg2 = _walk_by_expansion(counts, h2, order, L, skip)
        assert np.array_equal(idx, idx2)
        np.testing.assert_array_equal(ss, ss2)
        np.testing.assert_allclose(bg, bg2, rtol=1e-12)
    # per-frequency argmax (strict >, first in C order, n>0, h>0) and bg = sum - max
    ss, bg, idx, _ = oracle.ss_bg_counts(counts, h2)
    for f in range(F):
        for r in range(R):
            ok = (counts[:, f, r] > 0) & (h2[:, f] > 0)
            if ok.any():
                best = np.flatnonzero(ok)[np.argmax(h2[ok, f])]
                assert idx[f, r] == best
                assert ss[f, r] == h2[best, f]
            np.testing.assert_allclose(bg[f, r], np.sum(counts[:, f, r] * h2[:, f]) - ss[f, r], rtol=1e-12)
    gwb = oracle.poisson_gwb_counts(counts, h2)
    np.testing.assert_allclose(gwb, np.einsum('bfr,bf->fr', counts, h2), rtol=1e-12)


def test_dbn_gw_vs_reference_python_version():
    """Vectorised restatement of sam.py:618-676 (table comoving distance, ISCO by separation)."""
    from holodeck_amd import configs
    from oracle.cfg import oracle_cfg
    spec = configs.spec_config1(shape=(11, 9, 10), nfreq=6)
    cfg = oracle_cfg(spec)
    c = cfg['cosmo']
    cos = OracleCosmo(h=c['h'], Om0=c['Om0'])
    mtot, mrat, redz = oracle.edges_from_cfg(cfg)
    dens, zprime, _ = oracle.static_density(mtot, mrat, redz, cfg, cos)
    dens[zprime < 0] = 0.0
    fe = cfg['fobs_gw_edges']
    fo = oracle.midpoints(fe) / 2.0
    rf, dn = oracle.dbn_gw(fo, dens, mtot, mrat, zprime, cos)

    shp = dens.shape + (fo.size,)
    rz = zprime[..., None] * np.ones(shp)
    coal = rz > 0.0
    L = oracle.lib()
    cst = np.zeros(6)
    L.or_constants(cst.ctypes.data)
    gz, gd = cos._grid_z[::-1], cos._grid_dcom[::-1]
    dc = np.interp(rz[coal], gz, gd) / cst[5]
    fro = fo[None, None, None, :] * (1.0 + rz)
    mt = (mtot[:, None, None, None] * np.ones(shp))[coal]
    mr = (mrat[None, :, None, None] * np.ones(shp))[coal]
    fr = fro[coal]
    sa = cst[2] * mt ** (1 / 3) / fr ** (2 / 3)
    dadt = cst[0] * mt ** 3 * mr / sa ** 3 / (1 + mr) ** 2
    tau = -(2 / 3) * sa / dadt
    cosmo_fact = cst[3] * dc ** 2 * (1.0 + rz[coal])
    risco = 3.0 * cst[4] * mt
    live = sa > risco
    expect = np.zeros(shp)
    cc = coal.copy()
    cc[coal] = live
    expect[cc] = ((dens[..., None] * np.ones(shp))[coal] * tau * cosmo_fact)[live]
    np.testing.assert_allclose(dn, expect, rtol=1e-12, atol=0)
    assert np.all(rf[coal] == rz[coal]) and np.all(rf[~coal] == -1)


def test_dbn_2pwl_vs_continuous_integration():
    """redz_final and dnum at bracketed targets against quad() of dt = da/|da/dt| with exact cosmology."""
    from scipy.integrate import quad
    from holodeck_amd import configs
    from oracle.cfg import oracle_cfg
    spec = configs.spec_config2(shape=(7, 5, 6), nfreq=5)
    cfg = oracle_cfg(spec)
    o = oracle.number_h2fdf(cfg)
    cos, rf, dn = o['cosmo'], o['redz_final'], o['dnum']
    h = cfg['hard']
    L = oracle.lib()
    cst = np.zeros(6)
    L.or_constants(cst.ctypes.data)
    fo = oracle.midpoints(cfg['fobs_gw_edges']) / 2.0
    norm = 10.0 ** o['norm_log10']
    nchk = 0
    for i, mt in enumerate(o['mtot']):
        for j, mr in enumerate(o['mrat']):
            dadt = lambda a: L.or_hard_func_2pwl_gw(mt, mr, a, norm[i, j], h['rchar'], h['gamma_inner'],  # noqa
                                                    h['gamma_outer'])
            for k, z0 in enumerate(o['redz']):
                for f in range(fo.size):
                    zf = rf[i, j, k, f]
                    if zf <= 0.05:
                        continue
                    # separation where the binary is seen at fo: rest frame fo*(1+zf)
                    a_f = L.or_kepler_sepa(mt, fo[f] * (1 + zf))
                    lo, hi = np.log(a_f), np.log(h['sepa_init'])
                    t_evo = quad(lambda la: np.exp(la) / -dadt(np.exp(la)), lo, hi, limit=400, epsrel=1e-10)[0]
                    t_tot = cos.age(z0) + o['gmt_time'][i, j, k] + t_evo
                    z_exp = float(cos.tage_to_z(t_tot))
                    # 300-step trapezoid + 200-node table: agreement to ~1e-3 in 1+z
                    assert abs((1 + zf) / (1 + z_exp) - 1) < 3e-3, (i, j, k, f, zf, z_exp)
                    # dnum given the walk's own redshift: nden * tres * 4 pi c (1+z) dc^2
                    tres = -(2 / 3) * a_f / dadt(a_f)
                    dcz = np.interp(zf, cos._grid_z[::-1], cos._grid_dcom[::-1])
                    val = o['dens'][i, j, k] * tres * cst[3] * (1 + zf) * (dcz / cst[5]) ** 2
                    # the walk interpolates d_c in time, this check in z: table-level agreement
                    if o['dens'][i, j, k] == 0.0:
                        assert dn[i, j, k, f] == 0.0
                    else:
                        assert abs(dn[i, j, k, f] / val - 1) < 1e-3, (i, j, k, f)
                    nchk += 1
    assert nchk > 50
