"""Eccentric SAM GWB: the oracle pinned by the mathematics (CPU), and the HIP kernels against the
oracle (GPU).

Pins: the Peters-Mathews sum rule sum_n g(n, e) = F(e) for the direct formula, and the oracle
equal to that formula with the reference's -2 e J_n (cyutils.pyx:189) in place of -2 J_n; the
circular limit (only n = 2, h_c^2
proportional to f^-4/3); the harmonic ratios at constant eccentricity, g(n,e)/g(2,e) (2/n)^(2/3);
the Peters (1964) a(e) invariant along the RK4 evolution; and the discrete variant's Poisson
means summing (with their strain factors) to the deterministic spectrum.

Tolerance note: the reference's upward Bessel recurrence cancels at small e (relative error
about 2^-52 (2/e)^4), so g(n, e) at e below ~1e-4 carries rounding noise in the reference
itself; GPU-vs-oracle comparisons use evolutions that stay above that range in the band, and
the stated tolerance is 1e-6 per (frequency, harmonic) element plus 1e-12 of the harmonic sum,
1e-9 on the harmonic-summed spectrum.
"""
import numpy as np
import pytest

import oracle.eccen as OE

YR = 3.15576e7
PC = 3.085677581491367e18
MSOL = 1.988409870698051e33


def _grid(M=9, Q=7, Z=6, seed=1):
    mtot = np.logspace(np.log10(1e6 * MSOL), np.log10(1e10 * MSOL), M)
    mrat = np.logspace(-2, 0, Q)
    redz = np.logspace(-2, 0.5, Z)
    rng = np.random.default_rng(seed)
    ndens = 10 ** rng.uniform(-6, -2, (M, Q, Z))
    dcom = 4000.0 * redz / (1 + redz)        # [Mpc], any positive increasing values do
    fobs = np.arange(1, 6) / (16.03 * YR)
    return ndens, np.log10(mtot), mrat, redz, dcom, fobs, mtot


def _g_direct(n, e, c2):
    """g(n, e) from five scipy Bessel values, with coefficient c2 on J_n in the second bracket:
    c2 = 2 is Peters & Mathews (1963) Eq. 20; c2 = 2 e is the reference's cyutils.pyx:189."""
    import scipy.special as sps
    a, b, c, d, f = [sps.jv(n + k, n * e) for k in (-2, -1, 0, 1, 2)]
    aa = (a - 2 * e * b + (2 / n) * c + 2 * e * d - f) ** 2
    bb = (1 - e * e) * (a - c2 * c + f) ** 2
    return n ** 4 / 32 * (aa + bb + 4 / (3 * n * n) * c * c)


def test_g_sum_rule_and_the_reference_bracket():
    """With Peters & Mathews' second bracket, sum_n g(n, e) = F(e); the reference (and so the
    oracle) writes -2 e J_n there instead of -2 J_n, which breaks the sum rule at O(e^2): the
    oracle equals the direct formula with that coefficient, and its sum misses F(e)."""
    for e in (0.05, 0.2, 0.5, 0.7):
        n = range(1, 700)
        assert np.isclose(sum(_g_direct(k, e, 2.0) for k in n), OE.gw_ecc_func(e), rtol=1e-12), e
        ours = np.array([OE.gw_freq_dist_func(k, e) for k in range(1, 60)])
        direct = np.array([_g_direct(k, e, 2.0 * e) for k in range(1, 60)])
        assert np.allclose(ours, direct, rtol=1e-9, atol=1e-14 * OE.gw_ecc_func(e))
        assert not np.isclose(sum(OE.gw_freq_dist_func(k, e) for k in n), OE.gw_ecc_func(e), rtol=1e-4)
    assert OE.gw_freq_dist_func(2, 1e-13) == 1.0 and OE.gw_freq_dist_func(3, 1e-13) == 0.0


def test_evolution_keeps_the_peters_invariant():
    """a (1 - e^2) / (e^(12/19) (1 + 121/304 e^2)^(870/2299)) is constant along GW-driven evolution."""
    mtot = np.array([1e8 * MSOL, 1e9 * MSOL])
    sepa, ecc = OE.evolve_eccen_uniform_single(mtot, 0.6, 1.0 * PC, 4000, 2 * 6.6743e-8 / 29979245800.0 ** 2)
    ok = ecc > 1e-6
    c = sepa[ok] * (1 - ecc[ok] ** 2) / (ecc[ok] ** (12 / 19) * (1 + 121 / 304 * ecc[ok] ** 2) ** (870 / 2299))
    assert np.allclose(c, c[0], rtol=1e-6)
    assert ecc[0] == 0.6 and np.all(np.diff(ecc) <= 0)
    s0, e0 = OE.evolve_eccen_uniform_single(mtot, 0.0, 1.0 * PC, 50, 1.0)
    assert np.all(e0 == 0.0)


def _wide_window():
    sepa = np.logspace(25, 8, 60)            # [cm]: brackets every rest-frame frequency below
    return sepa


def test_circular_limit():
    """e = 0: only the n = 2 harmonic, and h_c^2 f^(4/3) is the same at every frequency."""
    ndens, ml, mr, rz, dc, fobs, _ = _grid()
    sepa = _wide_window()
    gwb = OE.sam_calc_gwb_single_eccen(ndens, ml, mr, rz, dc, fobs, sepa, np.zeros(sepa.size), 6)
    assert np.all(gwb[:, [0, 2, 3, 4, 5]] == 0.0) and np.all(gwb[:, 1] > 0)
    x = gwb[:, 1] * fobs ** (4 / 3)
    assert np.allclose(x, x[0], rtol=1e-12)


def test_harmonic_ratios_at_constant_eccentricity():
    """e fixed: gwb[f, n] / gwb[f, 2] = g(n,e)/g(2,e) (2/n)^(2/3) for every f."""
    ndens, ml, mr, rz, dc, fobs, _ = _grid()
    sepa = _wide_window()
    e = 0.4
    gwb = OE.sam_calc_gwb_single_eccen(ndens, ml, mr, rz, dc, fobs, sepa, np.full(sepa.size, e), 12)
    n = np.arange(1, 13)
    want = np.array([OE.gw_freq_dist_func(k, e) for k in n]) / OE.gw_freq_dist_func(2, e) * (2 / n) ** (2 / 3)
    assert np.allclose(gwb / gwb[:, 1:2], want[None, :], rtol=1e-10)


def test_discrete_means_sum_to_the_deterministic_spectrum():
    ndens, ml, mr, rz, dc, fobs, mtot = _grid()
    sepa, ecc = OE.evolve_eccen_uniform_single(mtot, 0.5, 1.0 * PC, 123, 2 * 6.6743e-8 / 29979245800.0 ** 2)
    det = OE.sam_calc_gwb_single_eccen(ndens, ml, mr, rz, dc, fobs, sepa, ecc, 10)
    lam, hv = OE.eccen_number_h(ndens, ml, mr, rz, dc, fobs, sepa, ecc, 10)
    assert np.allclose((lam * hv).sum(axis=(0, 1, 2)), det, rtol=1e-12, atol=0)
    assert (lam > 0).sum() > 0 and np.all(lam >= 0)


# ---------------------------------------------------------------------------
# GPU: the HIP kernels against the oracle
# ---------------------------------------------------------------------------
def _close(a, ref, rtol=1e-6, frac=1e-12):
    scale = np.abs(ref).sum(axis=-1, keepdims=True)
    return np.abs(a - ref) <= rtol * np.abs(ref) + frac * scale


@pytest.mark.gpu
def test_bessel_pair_matches_scipy(gpu):
    import scipy.special as sps
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print("[bessel pair]", flush=True)
    n = np.repeat(np.arange(1, 101), 14)
    e = np.tile([1e-12, 1e-9, 1e-6, 1e-4, 1e-3, 0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 0.95, 0.99, 0.999], 100)
    x = n * e
    lo = (n - 2).astype(np.int32)
    out = D.empty(2 * n.size)
    lo_d, x_d = D.to_dev(lo, dtype=torch.int32), D.to_dev(x)       # held until the kernel has run
    _lib.call("hd_bessel_jpair", D.ptr(lo_d), D.ptr(x_d), n.size, D.ptr(out), D.stream_ptr())
    got = D.to_host(out).reshape(-1, 2)
    ref = np.stack([sps.jv(lo, x), sps.jv(lo + 1, x)], axis=1)
    big = np.abs(ref) > 1e-250
    assert np.all(np.abs(got[big] - ref[big]) <= 1e-12 * np.abs(ref[big]))
    assert np.all(np.abs(got[~big]) <= 1e-240)


@pytest.mark.gpu
def test_g_matches_oracle(gpu):
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print("[g(n,e)]", flush=True)
    n = np.repeat(np.arange(1, 101), 9).astype(np.int32)
    e = np.tile([0.0, 1e-3, 0.01, 0.1, 0.3, 0.5, 0.7, 0.9, 0.99], 100)
    out = D.empty(n.size)
    n_d, e_d = D.to_dev(n, dtype=torch.int32), D.to_dev(e)         # held until the kernel has run
    _lib.call("hd_gw_freq_dist", D.ptr(n_d), D.ptr(e_d), n.size, D.ptr(out), D.stream_ptr())
    got = D.to_host(out)
    ref = np.array([OE.gw_freq_dist_func(int(a), float(b)) for a, b in zip(n, e)])
    Fe = np.array([OE.gw_ecc_func(b) for b in e])
    assert np.all(np.abs(got - ref) <= 1e-9 * np.abs(ref) + 1e-14 * Fe)


def _gpu_eccen_case(e0=0.5, nharms=24):
    ndens, ml, mr, rz, dc, fobs, mtot = _grid()
    sepa, ecc = OE.evolve_eccen_uniform_single(mtot, e0, 1.0 * PC, 123, 2 * 6.6743e-8 / 29979245800.0 ** 2)
    return (ndens, ml, mr, rz, dc, fobs, sepa, ecc), nharms


@pytest.mark.gpu
@pytest.mark.parametrize("e0", [0.0, 0.3, 0.7, 0.95])
def test_gwb_eccen_matches_oracle(e0, gpu):
    from holodeck_amd import cyutils
    print(f"[gwb eccen e0={e0}]", flush=True)
    args, H = _gpu_eccen_case(e0)
    ref = OE.sam_calc_gwb_single_eccen(*args, H)
    got = cyutils.sam_calc_gwb_single_eccen(*args, H)
    assert got.shape == ref.shape
    assert np.all(_close(got, ref))
    assert np.allclose(got.sum(axis=1), ref.sum(axis=1), rtol=1e-9, atol=0)


@pytest.mark.gpu
def test_eccen_number_matches_oracle(gpu):
    from holodeck_amd import engine
    from holodeck_amd import device as D
    print("[eccen number]", flush=True)
    args, H = _gpu_eccen_case(0.6, nharms=16)
    lam_r, h_r = OE.eccen_number_h(*args, H)
    M, Q, Z, F, _ = lam_r.shape
    dev = [D.to_dev(np.asarray(a, dtype=float)) for a in args]
    for h0, hc in ((0, 16), (5, 7)):
        lam, hv = engine.eccen_number(*dev, h0, hc)
        lam = D.to_host(lam).reshape(M, Q, Z, F, hc)
        hv = D.to_host(hv).reshape(M, Q, Z, F, hc)
        lr, hr = lam_r[..., h0:h0 + hc], h_r[..., h0:h0 + hc]
        both = (lam != 0) & (lr != 0)
        assert np.all(np.abs(lam[both] - lr[both]) <= 1e-9 * lr[both])
        # strain factors carry g(n, e): compare the contributions lambda * h, per frequency, to the
        # tolerance of the spectrum (1e-6 relative plus 1e-12 of the frequency's total)
        c, cr = lam * hv, lr * hr
        tot = (lam_r * h_r).sum(axis=(0, 1, 2, 4))[None, None, None, :, None]
        assert np.all(np.abs(c - cr) <= 1e-6 * np.abs(cr) + 1e-12 * tot)


@pytest.mark.gpu
def test_discrete_eccen_realizations_average_to_the_deterministic_spectrum(gpu):
    """Mean over realizations of the harmonic-summed spectrum equals the deterministic one within
    Monte-Carlo error, harmonic chunking changes nothing but the streams, and a seed repeats."""
    from holodeck_amd import cyutils
    print("[eccen discrete]", flush=True)
    args, H = _gpu_eccen_case(0.5, nharms=12)
    det = cyutils.sam_calc_gwb_single_eccen(*args, H).sum(axis=1)
    R = 4000
    g = cyutils.sam_calc_gwb_single_eccen_discrete(*args, H, R, seed=17, harm_chunk=5)
    assert g.shape == (5, H, R) and np.all(np.isfinite(g)) and np.all(g >= 0)
    s = g.sum(axis=1)                          # (F, R)
    z = (s.mean(axis=1) - det) / (s.std(axis=1) / np.sqrt(R))
    assert np.all(np.abs(z) < 5), z
    again = cyutils.sam_calc_gwb_single_eccen_discrete(*args, H, R, seed=17, harm_chunk=5)
    assert np.array_equal(again, g)


@pytest.mark.gpu
def test_eccentric_sam_surface(gpu):
    """gravwaves.sam_calc_gwb_single_eccen on a Semi_Analytic_Model with the evolution from
    sams.sam.evolve_eccen_uniform_single, against the oracle on the same density."""
    from holodeck_amd import gravwaves
    from holodeck_amd.sams.sam import Semi_Analytic_Model, evolve_eccen_uniform_single
    from holodeck_amd.constants import MPC
    print("[eccen SAM]", flush=True)
    sam = Semi_Analytic_Model(mtot=(1e6 * MSOL, 1e11 * MSOL, 13), mrat=(1e-2, 1.0, 9), redz=(1e-2, 3.0, 11))
    sepa, ecc = evolve_eccen_uniform_single(sam, 0.6, 1.0 * PC, 123)
    fobs = np.arange(1, 5) / (16.03 * YR)
    got = gravwaves.sam_calc_gwb_single_eccen(fobs, sam, sepa, ecc, nharms=20)
    dcom = sam.cosmo.comoving_distance(sam.redz) / MPC
    ref = OE.sam_calc_gwb_single_eccen(sam.static_binary_density, np.log10(sam.mtot), sam.mrat, sam.redz, dcom, fobs,
                                       sepa, ecc, 20)
    assert np.all(_close(got, ref)) and np.all(got.sum(axis=1) > 0)
    d = gravwaves.sam_calc_gwb_single_eccen_discrete(fobs, sam, sepa, ecc, nharms=20, nreals=8, seed=3)
    assert d.shape == (4, 20, 8) and np.all(np.isfinite(d))


def test_skipping_empty_cells_leaves_the_walk_unchanged():
    """The oracle walks only (mass, redshift) cells with some density; a grid whose empty cells
    hold a negligible density instead (so every cell is walked) gives the same spectrum."""
    ndens, ml, mr, rz, dc, fobs, mtot = _grid()
    rng = np.random.default_rng(3)
    empty = rng.uniform(size=(ndens.shape[0], ndens.shape[2])) < 0.6
    empty[0, :] = empty[:, 0] = True
    sparse = ndens * ~empty[:, None, :]
    tiny = sparse + 1e-280 * empty[:, None, :]
    sepa, ecc = OE.evolve_eccen_uniform_single(mtot, 0.6, 1.0 * PC, 123, 2 * 6.6743e-8 / 29979245800.0 ** 2)
    a = OE.sam_calc_gwb_single_eccen(sparse, ml, mr, rz, dc, fobs, sepa, ecc, 8)
    b = OE.sam_calc_gwb_single_eccen(tiny, ml, mr, rz, dc, fobs, sepa, ecc, 8)
    assert np.any(a > 0) and np.allclose(a, b, rtol=1e-13, atol=0)


# The eccentric library workload: 61x51x61 grid, 40 frequencies i/T, 100 harmonics.
def _config_61x51x61(e0):
    ndens, ml, mr, rz, dc, _, mtot = _grid(61, 51, 61, seed=7)
    fobs = np.arange(1, 41) / (16.03 * YR)
    sepa, ecc = OE.evolve_eccen_uniform_single(mtot, e0, 1.0 * PC, 123, 2 * 6.6743e-8 / 29979245800.0 ** 2)
    return ndens, ml, mr, rz, dc, fobs, sepa, ecc


def _cell_mask(M, Z, ncells=24, seed=5):
    """(M, Z) mask: the four corners (edge trapezoid weights) and random interior cells."""
    rng = np.random.default_rng(seed)
    mask = np.zeros((M, Z), bool)
    mask[[0, 0, M - 1, M - 1], [0, Z - 1, 0, Z - 1]] = True
    ii, kk = rng.integers(0, M, ncells - 4), rng.integers(0, Z, ncells - 4)
    mask[ii, kk] = True
    return mask


@pytest.mark.gpu
@pytest.mark.parametrize("e0", [0.3, 0.9])
def test_gwb_eccen_61x51x61_f40_h100_matches_oracle(e0, gpu):
    """Full eccentric-library shape against the oracle: the GPU spectrum of a density that is
    non-zero on 24 (mass, redshift) cells (all 51 mass ratios, 40 frequencies, 100 harmonics)
    equals the oracle's on the same grid; at the full density, the spectrum is additive over a
    split of the cells (the GPU sums independent cell terms)."""
    from holodeck_amd import cyutils
    print(f"[gwb eccen 61x51x61 e0={e0}]", flush=True)
    ndens, ml, mr, rz, dc, fobs, sepa, ecc = _config_61x51x61(e0)
    mask = _cell_mask(ndens.shape[0], ndens.shape[2])[:, None, :]
    sparse = ndens * mask
    ref = OE.sam_calc_gwb_single_eccen(sparse, ml, mr, rz, dc, fobs, sepa, ecc, 100)
    got = cyutils.sam_calc_gwb_single_eccen(sparse, ml, mr, rz, dc, fobs, sepa, ecc, 100)
    assert got.shape == (40, 100) and (ref > 0).sum() > 1000
    assert np.all(_close(got, ref))
    assert np.allclose(got.sum(axis=1), ref.sum(axis=1), rtol=1e-9, atol=0)
    full = cyutils.sam_calc_gwb_single_eccen(ndens, ml, mr, rz, dc, fobs, sepa, ecc, 100)
    rest = cyutils.sam_calc_gwb_single_eccen(ndens * ~mask, ml, mr, rz, dc, fobs, sepa, ecc, 100)
    scale = full.sum(axis=1, keepdims=True)
    assert np.all(np.abs(full - (got + rest)) <= 1e-12 * full + 1e-15 * scale)


@pytest.mark.gpu
def test_discrete_eccen_61x51x61_r200_mean(gpu):
    """200 realizations at the full eccentric-library shape (e0 = 0.5): per frequency, the
    Poisson means times strain factors sum to the deterministic spectrum, and the realization
    mean of the harmonic-summed spectrum lies within the exact standard error
    sqrt(sum h^2 lambda / R) of it -- |z| < 6 where the compound-Poisson skewness of the mean
    sum h^3 lambda / (sum h^2 lambda)^1.5 / sqrt(R) is below 1 (the lowest frequencies, where many
    binaries contribute), |z| < 30 (Chebyshev) elsewhere."""
    import torch
    from holodeck_amd import cyutils, engine
    from holodeck_amd import device as D
    print("[eccen discrete 61x51x61 R=200]", flush=True)
    args = _config_61x51x61(0.5)
    H, R, F = 100, 200, 40
    det = cyutils.sam_calc_gwb_single_eccen(*args, H).sum(axis=1)
    g = cyutils.sam_calc_gwb_single_eccen_discrete(*args, H, R, seed=11)
    assert g.shape == (F, H, R) and np.all(np.isfinite(g)) and np.all(g >= 0)
    dev = cyutils._eccen_dev(*args)
    k1, k2, k3 = (torch.zeros(F, dtype=torch.float64, device=dev[0].device) for _ in range(3))
    hc = 10
    for h0 in range(0, H, hc):
        lam, hv = engine.eccen_number(*dev, h0, hc)
        for k, p in ((k1, 1), (k2, 2), (k3, 3)):
            k += (lam * hv ** p).sum(dim=0).view(F, hc).sum(dim=1)
        del lam, hv
    k1, k2, k3 = (D.to_host(k) for k in (k1, k2, k3))
    assert np.allclose(k1, det, rtol=1e-9, atol=0)
    s = g.sum(axis=1)
    z = (s.mean(axis=1) - det) / np.sqrt(k2 / R)
    skew = k3 / k2 ** 1.5 / np.sqrt(R)
    near = skew < 1.0
    print("z", np.round(z, 2), "skew", np.round(skew, 3), flush=True)
    assert near.sum() >= 4
    assert np.all(np.abs(z[near]) < 6) and np.all(np.abs(z) < 30), (z, skew)
