"""Deterministic parity of the drawn realization path: the draw kernels record every count they draw
(hd_realize_record_counts), and the oracle's walks (cyutils.pyx:1266-1344, 1615-1767, 1068-1178,
862-897) run on exactly those counts.  Loud-source indices and strains must agree bit for bit; the
background sums and strain-weighted parameter means to 1e-12 of the total (the GPU sums per chunk
and merges, the oracle sums sequentially).

The full-size case runs the config-2 grid (91x81x101 edges: 720,000 bins, 40 frequencies) with
Fixed_Time_2PL_SAM, so the frequency-0 order differs from the per-frequency order, the draw kernel
has 88 chunks per frequency and the two-stage merge (groups of chunks, then all) runs.  The
synthetic case spans every sampler branch: sparse Poisson-process bins, table inversions, windows
too wide for the table (PTRS) and means above the Normal threshold.
"""
import numpy as np
import pytest
import torch

from holodeck_amd import configs

pytestmark = pytest.mark.gpu


def _rel_to_total(a, b, total, tol, what):
    err = np.abs(np.asarray(a) - np.asarray(b))
    bad = err > tol * np.abs(total)
    assert not bad.any(), f"{what}: {bad.sum()} outside {tol} of the total; worst {np.max(err / np.abs(total))}"


def _run(number, h2, R, L, mode, seed, order=None, flags=0, par=None, thresh=1e10):
    from holodeck_amd import engine
    from holodeck_amd import device as D
    B, F = h2.shape
    rec = torch.empty((B, F, R), dtype=torch.float64, device="cuda")
    out = engine.realize(number, h2, R, L, mode, order=order, flags=flags, par=par, seed=seed,
                         normal_threshold=thresh, record_counts=rec)
    plain = engine.realize(number, h2, R, L, mode, order=order, flags=flags, par=par, seed=seed,
                           normal_threshold=thresh)
    for a, b in zip(out, plain):   # recording changes nothing
        if a is not None:
            assert torch.equal(a, b)
    return [None if t is None else D.to_host(t) for t in out], D.to_host(rec)


def _check_sorted(oracle_lib, rec, h2, order, out, L, skip=False, par=None):
    ss, idx, bg, bgp = out
    ss_o, idx_o, bg_o, bgp_o = oracle_lib.loudest_counts(rec, h2, order, L, skip_zero_h=skip, par=par)
    assert np.array_equal(idx, idx_o), "loud indices"
    assert np.array_equal(ss, ss_o), "loud strains"
    total = bg_o + ss_o.sum(axis=-1)
    _rel_to_total(bg, bg_o, total, 1e-12, "hc2bg")
    if par is not None:
        # bgpar = sum(n h p) / sum(n h) over the background: compare the weighted sums, per parameter
        # relative to the total times the parameter's largest magnitude
        for k in range(par.shape[0]):
            _rel_to_total(bgp[k] * bg, bgp_o[k] * bg_o, total * np.max(np.abs(par[k])), 1e-12, f"bgpar[{k}]")
    return idx_o


@pytest.fixture(scope="module")
def full(gpu):
    print("[drawn parity: config-2 stage A]", flush=True)
    spec = configs.spec_config2()
    sam = configs.make_sam(spec)
    hard = configs.make_hard(spec, sam)
    number, h2fdf, rf = sam._number_h2fdf_device(configs.fobs_gw_edges(spec), hard)
    return spec, sam, number, h2fdf, rf


def test_drawn_sorted_walk_full_config2(full, oracle_lib):
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    spec, sam, number, h2fdf, _ = full
    R, L = 4, 10
    print(f"[drawn sorted walk B={number.shape[0]} F={number.shape[1]} R={R} L={L}]", flush=True)
    order = engine.order_desc(h2fdf, col=0)
    out, rec = _run(number, h2fdf, R, L, _lib.HD_SEL_SORTED, seed=11, order=order)
    h2 = D.to_host(h2fdf)
    od = D.to_host(order)
    assert np.array_equal(od, oracle_lib.order_desc(h2[:, 0]))
    lam = D.to_host(number)
    assert np.all(rec[lam == 0.0] == 0.0)
    idx = _check_sorted(oracle_lib, rec, h2, od, out, L)
    # the loud lists at f > 0 follow the frequency-0 order, not the per-frequency strain order
    assert np.any(idx[1:, :, 1:] >= 0)
    # the injected-count kernel on the same counts (its own chunking, the same two-stage merge)
    ss2, idx2, bg2, _ = engine.realize(number, h2fdf, R, L, _lib.HD_SEL_SORTED, order=order,
                                       counts=D.to_dev(rec))
    assert np.array_equal(D.to_host(idx2), out[1])
    total = out[2] + out[0].sum(axis=-1)
    _rel_to_total(D.to_host(bg2), out[2], total, 1e-12, "injected vs drawn hc2bg")
    del rec


def test_drawn_sorted_walk_with_params_full_config2(full, oracle_lib):
    """loudest_hc_and_par_from_sorted_redz semantics: zero-strain bins skipped, 7 parameters."""
    from holodeck_amd import engine, _lib, single_sources
    from holodeck_amd import device as D
    spec, sam, number, h2fdf, rf = full
    R, L = 2, 5
    print(f"[drawn sorted walk with params R={R} L={L}]", flush=True)
    edges = [sam.mtot, sam.mrat, sam.redz, configs.fobs_gw_edges(spec) / 2.0]
    par = single_sources.bin_params_redz(edges, rf, sam.cosmo)
    order = engine.order_desc(h2fdf, col=0)
    out, rec = _run(number, h2fdf, R, L, _lib.HD_SEL_SORTED, seed=12, order=order,
                    flags=_lib.HD_FLAG_SKIP_ZERO_H, par=par)
    p = D.to_host(par)
    # parameters with infinite values (d_c where z_final <= 0) only enter bins with zero strain
    fin = np.where(np.isfinite(p), p, 0.0)
    _check_sorted(oracle_lib, rec, D.to_host(h2fdf), D.to_host(order), out, L, skip=True, par=fin)


def test_drawn_argmax_and_none_full_config2(full, oracle_lib):
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    spec, sam, number, h2fdf, _ = full
    R = 3
    print(f"[drawn argmax/none R={R}]", flush=True)
    B, F = h2fdf.shape
    rng = np.random.default_rng(4)
    par_b = rng.uniform(1.0, 2.0, size=(3, B))
    par = D.to_dev(np.repeat(par_b[:, :, None], F, axis=2))
    out, rec = _run(number, h2fdf, R, 1, _lib.HD_SEL_ARGMAX, seed=13, par=par)
    h2 = D.to_host(h2fdf)
    ss_o, bg_o, idx_o, bgp_o = oracle_lib.ss_bg_counts(rec, h2, par=par_b)
    assert np.array_equal(out[1][..., 0], idx_o)
    assert np.array_equal(out[0][..., 0], ss_o)
    total = bg_o + ss_o
    _rel_to_total(out[2], bg_o, total, 1e-12, "argmax hc2bg")
    _rel_to_total(out[3] * out[2], bgp_o * bg_o, 2.0 * total, 1e-12, "argmax bgpar")
    del rec
    out, rec = _run(number, h2fdf, R, 0, _lib.HD_SEL_NONE, seed=14)
    gwb_o = oracle_lib.poisson_gwb_counts(rec, h2)
    _rel_to_total(out[2], gwb_o, gwb_o, 1e-12, "poisson gwb")


def test_drawn_every_sampler_branch(gpu, oracle_lib):
    """Means from 1e-6 to 1e12 (sparse bins, CDF tables, PTRS for windows wider than the table,
    Normal above 1e10), zeros and zero strains; three modes against the oracle on the recorded
    counts, and the recorded counts' moments against the means."""
    from holodeck_amd import engine, _lib
    from holodeck_amd import device as D
    print("[drawn every sampler branch]", flush=True)
    rng = np.random.default_rng(21)
    B, F, R, L = 30011, 3, 96, 4
    lam = 10.0 ** rng.uniform(-6, 12, size=(B, F))
    lam[rng.random((B, F)) < 0.05] = 0.0
    h2 = 10.0 ** rng.uniform(-35, -25, size=(B, F))
    h2[rng.random((B, F)) < 0.03] = 0.0
    order = oracle_lib.order_desc(h2[:, 0])
    n_d, h_d, o_d = D.to_dev(lam), D.to_dev(h2), D.to_dev(order)
    for skip in (False, True):
        out, rec = _run(n_d, h_d, R, L, _lib.HD_SEL_SORTED, seed=31 + skip, order=o_d,
                        flags=_lib.HD_FLAG_SKIP_ZERO_H if skip else 0)
        _check_sorted(oracle_lib, rec, h2, order, out, L, skip=skip)
    # integer counts below the Normal threshold; moments by regime (the last run skipped zero strains)
    drawn = (lam > 0) & (h2 != 0.0)
    assert np.all(rec[~drawn] == 0.0)
    small = drawn & (lam <= 1e10)
    assert np.all(rec[small] == np.round(rec[small]))
    for lo, hi in ((1e-6, 1e-2), (1e-2, 1.0), (1.0, 1e3), (1e3, 1e6), (1e6, 1e10), (1e10, 1e12)):
        sel = drawn & (lam > lo) & (lam <= hi)
        tot = lam[sel].sum() * R          # the summed count is Poisson (Normal) with this mean
        assert abs(rec[sel].sum() - tot) < 5.0 * np.sqrt(tot), (lo, hi)
        if lo >= 1.0:                     # per-bin standardised means: unit variance
            z = (rec[sel].mean(axis=-1) - lam[sel]) / np.sqrt(lam[sel] / R)
            assert 0.85 < z.var() < 1.15, (lo, hi, z.var())
    out, rec = _run(n_d, h_d, R, 1, _lib.HD_SEL_ARGMAX, seed=33)
    ss_o, bg_o, idx_o, _ = oracle_lib.ss_bg_counts(rec, h2)
    assert np.array_equal(out[1][..., 0], idx_o)
    _rel_to_total(out[2], bg_o, bg_o + ss_o, 1e-12, "argmax hc2bg")
