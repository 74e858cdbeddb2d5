"""Statistics of the on-GPU counter-based realizations.

* the per-cell sampler against the Poisson (and, above 1e10, Normal) law:
  moments and a chi-square / KS test per lambda regime;
* the realized h_c (background and loudest) against the CPU oracle's own
  independent generator (xoshiro256** + the numpy-documented samplers):
  per-frequency two-sample KS p > 0.01 (family-wise over a test's comparisons) and
  median / 5th / 95th percentiles within order-statistic bounds.  The reference's draws are unseeded
  (cyutils.pyx:1313-1317), so agreement is distributional.
"""
import numpy as np
import pytest
from scipy import stats

from holodeck_amd import configs

pytestmark = pytest.mark.gpu


def _draws(lam, R, seed):
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    n = D.to_dev(np.array([[lam]]))
    h = D.to_dev(np.array([[1.0]]))
    _, _, bg, _ = engine.realize(n, h, R, 0, _lib.HD_SEL_NONE, seed=seed)
    return D.to_host(bg)[0]


@pytest.mark.parametrize("lam", [1e-7, 0.02, 0.7, 3.0, 9.99, 10.0, 47.5, 1.0e3, 3.3e6])
def test_poisson_sampler(lam, gpu):
    print(f"[poisson lam={lam}]", flush=True)
    R = 200000
    x = _draws(lam, R, seed=17)
    assert np.all(x == np.floor(x)) and np.all(x >= 0)
    se = np.sqrt(lam / R)
    assert abs(x.mean() - lam) < 5 * se + 1e-12
    if lam > 1e-3:
        assert abs(x.var() / lam - 1) < 5 * np.sqrt(2.0 / R) + 5 / np.sqrt(R * lam)
    # chi-square over the bulk of the pmf
    kmax = int(stats.poisson.ppf(1 - 1e-4, lam))
    kmin = int(stats.poisson.ppf(1e-4, lam))
    ks = np.arange(kmin, kmax + 1)
    obs = np.array([np.sum(x == k) for k in ks], dtype=float)
    exp = stats.poisson.pmf(ks, lam) * R
    ok = exp > 20
    if ok.sum() >= 2:
        o, e = obs[ok], exp[ok]
        chi2 = np.sum((o - e) ** 2 / e)
        p = stats.chi2.sf(chi2, ok.sum() - 1)
        assert p > 1e-3, (lam, chi2, ok.sum(), p)
    if lam < 1e-5:
        assert np.sum(x > 0) <= 5


def test_per_bin_laws_inside_full_tiles(gpu):
    """Every frequency holds 3000 bins with lambda log-uniform over 1e-4..3e5 (sparse bins, table
    windows of every size, batches split by the LDS table, bins too wide for it); the strain is 1 on
    one chosen bin and 0 elsewhere, so hc2bg[f] is that bin's own count.  Chi-square per frequency."""
    print("[per-bin laws in full tiles]", flush=True)
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    rng = np.random.default_rng(11)
    B, F, R = 3000, 48, 4000
    lam = 10.0 ** rng.uniform(-4, np.log10(3e5), size=(B, F))
    sel = rng.integers(0, B, size=F)
    # pin a spread of regimes on the selected bins
    pinned = [3e-4, 0.05, 0.6, 0.999, 1.0, 2.5, 9.9, 10.0, 33.0, 120.0, 900.0, 4.2e3, 2.0e4, 9.0e4, 2.9e5]
    for f, v in enumerate(pinned):
        lam[sel[f], f] = v
    h = np.zeros((B, F))
    h[sel, np.arange(F)] = 1.0
    _, _, bg, _ = engine.realize(D.to_dev(lam), D.to_dev(h), R, 0, _lib.HD_SEL_NONE, seed=77)
    x = D.to_host(bg)
    pvals = []
    for f in range(F):
        l = lam[sel[f], f]
        xf = x[f]
        assert np.all(xf == np.floor(xf)) and np.all(xf >= 0), f
        assert abs(xf.mean() - l) < 6 * np.sqrt(l / R) + 1e-12, (f, l, xf.mean())
        kmin, kmax = int(stats.poisson.ppf(1e-4, l)), int(stats.poisson.ppf(1 - 1e-4, l))
        ks = np.arange(kmin, kmax + 1)
        e = stats.poisson.pmf(ks, l) * R
        ok = e > 20
        if ok.sum() >= 2:
            o = np.array([np.sum(xf == k) for k in ks], dtype=float)[ok]
            chi2 = np.sum((o - e[ok]) ** 2 / e[ok])
            pvals.append(stats.chi2.sf(chi2, ok.sum() - 1))
    assert min(pvals) > 1e-3 / len(pvals), sorted(pvals)[:5]
    # the background of all bins at once is Poisson(sum lambda) when h = 1 everywhere
    _, _, bg1, _ = engine.realize(D.to_dev(lam[:, :4]), D.to_dev(np.ones((B, 4))), R, 0, _lib.HD_SEL_NONE, seed=78)
    tot = D.to_host(bg1)
    for f in range(4):
        S = lam[:, f].sum()
        assert abs(tot[f].mean() - S) < 6 * np.sqrt(S / R), (f, S, tot[f].mean())
        assert abs(tot[f].var() / S - 1) < 6 * np.sqrt(2.0 / R)


def test_normal_branch_above_threshold(gpu):
    lam = 3.0e10
    x = _draws(lam, 100000, seed=5)
    z = (x - lam) / np.sqrt(lam)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    assert stats.kstest(z, 'norm').pvalue > 1e-3
    assert np.any(x != np.floor(x))   # not floored, as the reference


def test_streams_independent_of_launch_and_seeded(gpu):
    a = _draws(2.5, 5000, seed=3)
    b = _draws(2.5, 5000, seed=3)
    c = _draws(2.5, 5000, seed=4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    # a prefix of the realizations does not depend on R (counter keyed by r)
    d = _draws(2.5, 777, seed=3)
    assert np.array_equal(a[:777], d)


def _percentiles_within_order_statistics(a, b, what, m=1):
    """The 5th, 50th and 95th percentiles of sample a against sample b by order statistics: b's
    empirical CDF at a's q-th percentile must reach q within z standard errors of the difference
    of two independent empirical CDFs there, sqrt(q (1 - q) (1/Ra + 1/Rb)); equivalently, a's
    percentile lies between b's order statistics of ranks Rb (q -/+ z se).  z is two-sided 1%
    over the 3 m checks of a test (Bonferroni)."""
    z = stats.norm.isf(0.005 / (3 * m))
    for q in (0.05, 0.5, 0.95):
        qa = np.quantile(a, q, method="inverted_cdf")
        # b's empirical CDF just below and at qa (they differ where qa is an atom: loud-source
        # strains take the values of single bins)
        lo, hi = np.mean(b < qa), np.mean(b <= qa)
        tol = z * np.sqrt(q * (1.0 - q) * (1.0 / a.size + 1.0 / b.size)) + 1.0 / b.size
        assert lo - tol <= q <= hi + tol, (what, q, lo, hi, tol)


def _ks(a, b, what, m=1):
    """Two-sample KS, p > 0.01 over the m comparisons of a test (Bonferroni: each p > 0.01 / m)."""
    p = stats.ks_2samp(a, b).pvalue
    assert p > 0.01 / m, (what, p, m)


def test_hc_distribution_vs_cpu_oracle(gpu, oracle_lib):
    """Realized hc_bg and loudest hc_ss: GPU generator vs the oracle's generator, same number/h2fdf
    (GW-only grid)."""
    print("[hc distribution vs oracle]", flush=True)
    from oracle.cfg import oracle_cfg
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    spec = configs.spec_config1(shape=(21, 17, 15), nfreq=5)
    o = oracle_lib.number_h2fdf(oracle_cfg(spec))
    F = o['number'].shape[-1]
    n2 = o['number'].reshape(-1, F)
    h2 = o['h2fdf'].reshape(-1, F)
    order = oracle_lib.order_desc(h2[:, 0])
    R, L = 1000, 3
    ss_c, bg_c = oracle_lib.loudest_rng(n2, h2, order, R, L, seed=99)
    ss_g, _, bg_g, _ = engine.realize(D.to_dev(n2), D.to_dev(h2), R, L, _lib.HD_SEL_SORTED, order=D.to_dev(order),
                                      seed=2024)
    ss_g, bg_g = D.to_host(ss_g), D.to_host(bg_g)
    for f in range(F):
        for name, a, b in (("bg", bg_g[f], bg_c[f]), ("ss0", ss_g[f, :, 0], ss_c[f, :, 0])):
            _ks(a, b, (name, f), m=2 * F)
            _percentiles_within_order_statistics(a, b, (name, f), m=2 * F)


def test_hc_and_params_distribution_2pl_vs_cpu_oracle(gpu, oracle_lib):
    """Fixed_Time_2PL_SAM grid (the frequency-0 order differs from the later frequencies' order),
    loudest_hc_and_par_from_sorted_redz semantics (zero strains skipped, 7 parameters): at every
    frequency the GPU's hc_bg, the three loudest hc_ss and the strain-weighted background
    parameters over 1000 realizations against the oracle's own generator on the same inputs
    (two-sample KS p > 0.01; 5th / 50th / 95th percentiles within order-statistic bounds)."""
    print("[hc + params distribution, 2PL, vs oracle]", flush=True)
    from oracle.cfg import oracle_cfg
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    spec = configs.spec_config2(shape=(21, 17, 15), nfreq=5)
    o = oracle_lib.number_h2fdf(oracle_cfg(spec))
    F = o['number'].shape[-1]
    n2 = o['number'].reshape(-1, F)
    h2 = o['h2fdf'].reshape(-1, F)
    par = oracle_lib.bin_params_redz(o['edges'], o['redz_final'], o['cosmo'])
    par = np.where(np.isfinite(par), par, 0.0)      # non-finite only where the strain is zero (skipped)
    order = oracle_lib.order_desc(h2[:, 0])
    R, L = 1000, 3
    counts = oracle_lib.draw_counts(n2, R, seed=77)
    ss_c, _, bg_c, bp_c = oracle_lib.loudest_counts(counts, h2, order, L, skip_zero_h=True, par=par)
    del counts
    ss_g, _, bg_g, bp_g = engine.realize(D.to_dev(n2), D.to_dev(h2), R, L, _lib.HD_SEL_SORTED,
                                         order=D.to_dev(order), flags=_lib.HD_FLAG_SKIP_ZERO_H,
                                         par=D.to_dev(par), seed=4048)
    ss_g, bg_g, bp_g = D.to_host(ss_g), D.to_host(bg_g), D.to_host(bp_g)
    ks = (0, 1, 2, 3, 5)
    m = F * (1 + L + len(ks))
    for f in range(F):
        pairs = [("bg", bg_g[f], bg_c[f])] + [(f"ss{j}", ss_g[f, :, j], ss_c[f, :, j]) for j in range(L)]
        pairs += [(f"bgpar{k}", bp_g[k, f], bp_c[k, f]) for k in ks]
        for name, a, b in pairs:
            _ks(a, b, (name, f), m=m)
            _percentiles_within_order_statistics(a, b, (name, f), m=m)
