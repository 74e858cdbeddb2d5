"""The full-size workload (91x81x101 edges, 40 frequencies, 500 realizations, 10 loudest) through
size-independent properties: every loud source is a real bin's strain, the frequency-0 loud lists
follow the sorted walk, loud + background equal the undivided total drawn from the same streams,
the realization mean matches the expectation sum(number * h2fdf) within Monte-Carlo error, and a
seed repeats bit for bit."""
import numpy as np
import pytest


@pytest.mark.gpu
def test_full_size_gwb_properties(gpu):
    import torch
    from holodeck_amd import configs, engine, _lib
    from holodeck_amd.cosmology import Cosmology
    from holodeck_amd import device as D
    print("[full size 91x81x101 F=40 R=500 L=10]", flush=True)
    spec = configs.spec_config2()
    fe = configs.fobs_gw_edges(spec)
    cos = Cosmology(h=spec['cosmo']['h'], Om0=spec['cosmo']['Om0'])
    sam = configs.make_sam(spec, cosmo=cos)
    hard = configs.make_hard(spec, sam)
    number, h2fdf, _ = sam._number_h2fdf_device(fe, hard)
    B, F = h2fdf.shape
    R, L = 500, 10
    order = engine.order_desc(h2fdf, col=0)
    ss, idx, bg, _ = engine.realize(number, h2fdf, R, L, _lib.HD_SEL_SORTED, order=order, seed=21)
    _, _, tot, _ = engine.realize(number, h2fdf, R, 0, _lib.HD_SEL_NONE, seed=21)
    ss, idx, bg, tot = (D.to_host(t) for t in (ss, idx, bg, tot))
    h = D.to_host(h2fdf)
    assert ss.shape == (F, R, L) and np.all(np.isfinite(bg)) and np.all(bg > 0)
    # every selected source is a bin's strain at its frequency
    ok = idx >= 0
    assert ok.mean() > 0.99 and np.all(idx[ok] < B)
    f_of = np.broadcast_to(np.arange(F)[:, None, None], idx.shape)
    assert np.array_equal(ss[ok], h[idx[ok], f_of[ok]])
    # frequency 0 walks bins in descending h2fdf: its loud lists are non-increasing
    assert np.all(np.diff(ss[0], axis=-1) <= 0)
    # loud + background = the total of the same draws (counts are integers below 1e10)
    assert np.allclose(bg + ss.sum(axis=-1), tot, rtol=1e-12, atol=0)
    # the mean over realizations is the expectation sum(lambda h), to within 5 standard errors of
    # the exact variance sum(lambda h^2) / R.  That variance is carried by rare loud bins
    # (lambda < 0.01), so the bound is loose at high frequencies; at frequency 0 it still catches a
    # 5% error in the mean.
    lam_h = number * h2fdf
    expect = D.to_host(lam_h.sum(dim=0))
    var = D.to_host((lam_h * h2fdf).sum(dim=0))
    z = (tot.mean(axis=1) - expect) / np.sqrt(var / R)
    assert np.all(np.abs(z) < 5), z
    assert abs(tot[0].mean() / expect[0] - 1.0) < 0.05
    # a seed repeats bit for bit
    ss2, idx2, bg2, _ = engine.realize(number, h2fdf, R, L, _lib.HD_SEL_SORTED, order=order, seed=21)
    assert np.array_equal(D.to_host(ss2), ss) and np.array_equal(D.to_host(bg2), bg)
    torch.cuda.synchronize()
