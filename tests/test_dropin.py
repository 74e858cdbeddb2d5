"""The Python drop-ins driven the way the reference's own modules would drive them: with
reference-shaped objects that are not this package's classes (plain attribute holders whose class
names are the reference's), and with the caller's cosmology grids (the oracle's, not this
package's), against the oracle (sam_cyutils.pyx:401-504, cyutils.pyx:1347-1538)."""
import numpy as np
import pytest

from holodeck_amd import configs

pytestmark = pytest.mark.gpu


class Fixed_Time_2PL_SAM:          # the reference's class name; attributes as hardening.py:1377-1423 sets them
    pass


class Hard_GW:
    pass


class _SAM:                         # the attributes sam_cyutils.pyx:457-494 reads
    pass


def _ref_shaped(spec, o, with_gmt=True):
    sam = _SAM()
    sam.static_binary_density = o['dens']
    sam.shape = o['dens'].shape
    sam.mtot, sam.mrat, sam.redz = o['mtot'], o['mrat'], o['redz']
    sam._gmt_time = o['gmt_time'] if with_gmt else None
    sam._redz_prime = o['zprime'] if with_gmt else None
    return sam


def _close(a, b, rtol, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    err = np.abs(a - b)
    assert np.all(err <= rtol * np.abs(b)), f"{what}: worst {np.max(err / (np.abs(b) + 1e-300))}"


def test_dbn_2pwl_with_reference_shaped_objects_and_caller_grids(gpu, oracle_lib):
    from oracle.cfg import oracle_cfg
    from holodeck_amd.sams import sam_cyutils
    print("[drop-in dbn 2pwl]", flush=True)
    spec = configs.spec_config2(shape=(13, 11, 9), nfreq=6)
    cfg = oracle_cfg(spec)
    cfg['mmbulge']['scatter_dex'] = 0.0
    o = oracle_lib.number_h2fdf(cfg)
    h = cfg['hard']
    hard = Fixed_Time_2PL_SAM()
    hard._sepa_init, hard._num_steps, hard._rchar = h['sepa_init'], h['num_steps'], h['rchar']
    hard._gamma_inner, hard._gamma_outer = h['gamma_inner'], h['gamma_outer']
    hard._norm = 10.0 ** o['norm_log10']
    fo = oracle_lib.midpoints(cfg['fobs_gw_edges']) / 2.0
    # the oracle's cosmology on a 137-node grid: the drop-in must use the caller's tables
    cos = oracle_lib.OracleCosmo(h=cfg['cosmo']['h'], Om0=cfg['cosmo']['Om0'], size=137, zmax=15.0)
    for with_gmt in (True, False):
        sam = _ref_shaped(spec, o, with_gmt)
        rf, dn = sam_cyutils.dynamic_binary_number_at_fobs(fo, sam, hard, cos)
        gt = o['gmt_time'] if with_gmt else np.zeros(o['dens'].shape)
        rf_o, dn_o = oracle_lib.dbn_2pwl(fo, h['sepa_init'], h['num_steps'], hard._norm, h['rchar'], h['gamma_inner'],
                                         h['gamma_outer'], o['dens'], o['mtot'], o['mrat'], o['redz'], gt, cos)
        _close(rf, rf_o, 1e-9, "redz_final")
        _close(dn, dn_o, 1e-9, "diff_num")


def test_dbn_gw_with_reference_shaped_objects(gpu, oracle_lib):
    from oracle.cfg import oracle_cfg
    from holodeck_amd.sams import sam_cyutils
    print("[drop-in dbn gw]", flush=True)
    spec = configs.spec_config1(shape=(13, 11, 9), nfreq=6)
    cfg = oracle_cfg(spec)
    cfg['mmbulge']['scatter_dex'] = 0.0
    o = oracle_lib.number_h2fdf(cfg)
    fo = oracle_lib.midpoints(cfg['fobs_gw_edges']) / 2.0
    cos = oracle_lib.OracleCosmo(h=cfg['cosmo']['h'], Om0=cfg['cosmo']['Om0'], size=211, zmax=12.0)
    for hard in (Hard_GW(), Hard_GW):                 # an instance or the class itself, as the reference allows
        for with_gmt in (True, False):
            sam = _ref_shaped(spec, o, with_gmt)
            rf, dn = sam_cyutils.dynamic_binary_number_at_fobs(fo, sam, hard, cos)
            rp = o['zprime'] if with_gmt else np.broadcast_to(o['redz'][None, None, :], o['dens'].shape)
            rf_o, dn_o = oracle_lib.dbn_gw(fo, o['dens'], o['mtot'], o['mrat'], rp, cos)
            _close(rf, rf_o, 1e-12, "redz_final")
            _close(dn, dn_o, 1e-12, "diff_num")
    with pytest.raises(ValueError, match="Unexpected"):
        sam_cyutils.dynamic_binary_number_at_fobs(fo, _ref_shaped(spec, o), object(), cos)


def test_integrate_binary_evolution_2pwl_matches_oracle(gpu, oracle_lib):
    from holodeck_amd.sams import sam_cyutils
    from holodeck_amd.constants import MSOL, PC
    print("[drop-in lifetime 2pwl]", flush=True)
    L = oracle_lib.lib()
    rng = np.random.default_rng(2)
    nl = rng.uniform(-8, 2, 40)
    mt = 10.0 ** rng.uniform(4, 12, 40) * MSOL
    mr = 10.0 ** rng.uniform(-3, 0, 40)
    args = (1e4 * PC, 100 * PC, -1.0, 2.5, 300)
    got = sam_cyutils.integrate_binary_evolution_2pwl(nl, mt, mr, *args)
    want = np.array([L.or_lifetime_2pwl(a, b, c, 0.0, *args) for a, b, c in zip(nl, mt, mr)])
    _close(got, want, 1e-12, "lifetime")
    one = sam_cyutils.integrate_binary_evolution_2pwl(nl[0], mt[0], mr[0], *args)
    assert isinstance(one, float) and abs(one / want[0] - 1) < 1e-12


def test_loudest_hc_and_par_from_sorted_with_injected_counts(gpu, oracle_lib):
    from holodeck_amd import cyutils
    print("[drop-in loudest_hc_and_par_from_sorted]", flush=True)
    rng = np.random.default_rng(8)
    M, Q, Z, F, R, L = 9, 8, 7, 4, 6, 3
    h2 = 10.0 ** rng.uniform(-35, -25, size=(M, Q, Z, F))
    h2[rng.random(h2.shape) < 0.05] = 0.0           # zero strains are kept by this variant
    number = 10.0 ** rng.uniform(-3, 0.7, size=(M, Q, Z, F))
    counts = rng.poisson(number[..., None] * np.ones(R)).astype(float)
    mt, mr, rz = rng.uniform(1, 2, M), rng.uniform(0.1, 1, Q), rng.uniform(0.01, 3, Z)
    flat_order = oracle_lib.order_desc(h2.reshape(-1, F)[:, 0])
    ms, qs, zs = np.unravel_index(flat_order, (M, Q, Z))
    ss, bg, lspar, bgpar, idx = cyutils.loudest_hc_and_par_from_sorted(number, h2, R, L, mt, mr, rz, ms, qs, zs,
                                                                       counts=counts)
    mm, qq, zz = np.meshgrid(np.arange(M), np.arange(Q), np.arange(Z), indexing='ij')
    par = np.stack([np.repeat(p.reshape(-1, 1), F, axis=1) for p in (mt[mm], mr[qq], rz[zz])])
    ss_o, idx_o, bg_o, bgp_o = oracle_lib.loudest_counts(counts.reshape(-1, F, R), h2.reshape(-1, F), flat_order, L,
                                                         par=par)
    assert np.array_equal(ss, ss_o)
    ok = idx_o >= 0
    m3 = np.unravel_index(np.where(ok, idx_o, 0), (M, Q, Z))
    for k in range(3):
        assert np.array_equal(idx[k], np.where(ok, m3[k], 0))
    _close(bg, bg_o, 1e-12, "hc2bg")
    _close(bgpar, bgp_o, 1e-12, "bgpar")
    vals = (mt[m3[0]], mr[m3[1]], rz[m3[2]])
    with np.errstate(invalid='ignore'):
        ls_o = np.stack([np.where(ok, ss_o * v, 0).sum(-1) / ss_o.sum(-1) for v in vals])
    assert np.array_equal(np.isnan(lspar), np.isnan(ls_o))
    fin = ~np.isnan(ls_o)
    _close(lspar[fin], ls_o[fin], 1e-12, "lspar")
