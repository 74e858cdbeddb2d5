"""hd_realize_noalloc (the realization launch sequence on the library's fixed device scratch array)
against hd_realize (the same kernels on a caller workspace), through the C-ABI: for the same seed
every output must be identical, on two calls in a row, the second on a smaller input (the scratch
then still holds the first call's contents).  Sizes beyond the fixed scratch are refused on the
host with HD_ERR_SIZE, before anything is launched."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

HD_ERR_SIZE = -3
SCRATCH_BYTES = 256 * 1024 * 1024   # HD_NOALLOC_SCRATCH_BYTES in include/holohip.h


def _inputs(rng, B, F, npar):
    """Means over every sampler branch (sparse, CDF tables, windows wider than the table, Normal
    above 1e10), zeros, and strains with zeros."""
    lam = 10.0 ** rng.uniform(-6, 12, size=(B, F))
    lam[rng.random((B, F)) < 0.05] = 0.0
    h2 = 10.0 ** rng.uniform(-35, -25, size=(B, F))
    h2[rng.random((B, F)) < 0.03] = 0.0
    par = rng.uniform(1.0, 2.0, size=(npar, B, F)) if npar else None
    return lam, h2, par


def _noalloc(number, h2fdf, order, R, L, mode, flags, seed, par):
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    B, F = h2fdf.shape
    Lk = 0 if mode == _lib.HD_SEL_NONE else L
    npar = 0 if par is None else int(par.shape[0])
    # outputs poisoned, so that every element must be written
    hc2bg = torch.full((F, R), float("nan"), dtype=torch.float64, device="cuda")
    hc2ss = torch.full((F, R, max(Lk, 1)), float("nan"), dtype=torch.float64, device="cuda")
    ssidx = torch.full((F, R, max(Lk, 1)), -7, dtype=torch.int64, device="cuda")
    bgpar = torch.full((npar, F, R), float("nan"), dtype=torch.float64, device="cuda") if npar else None
    _lib.call("hd_realize_noalloc", D.ptr(number), D.ptr(h2fdf), D.ptr(order), B, F, R, Lk, mode, flags, 1e10,
              seed, D.ptr(par), npar, D.ptr(hc2ss), D.ptr(ssidx), D.ptr(hc2bg), D.ptr(bgpar), D.stream_ptr())
    if Lk == 0:
        hc2ss, ssidx = hc2ss[..., :0], ssidx[..., :0]
    return hc2ss, ssidx, hc2bg, bgpar


CASES = [   # (B, F, R, L, mode, flags, npar, seed): the second is smaller than the first
    (200003, 6, 64, 5, 0, 0, 0, 101),
    (30011, 3, 96, 1, 1, 0, 3, 102),
]


def test_noalloc_matches_hd_realize_two_calls_in_a_row(gpu):
    from holodeck_amd import _lib, engine
    from holodeck_amd import device as D
    rng = np.random.default_rng(5)
    for B, F, R, L, mode, flags, npar, seed in CASES:
        print(f"[no-alloc realize B={B} F={F} R={R} L={L} mode={mode} npar={npar}]", flush=True)
        assert _lib.size("hd_realize_workspace_size", B, F, R, L, npar) <= SCRATCH_BYTES
        lam, h2, par = _inputs(rng, B, F, npar)
        n_d, h_d = D.to_dev(lam), D.to_dev(h2)
        p_d = D.to_dev(par) if npar else None
        order = engine.order_desc(h_d, col=0) if mode == _lib.HD_SEL_SORTED else None
        ref = engine.realize(n_d, h_d, R, L, mode, order=order, flags=flags, par=p_d, seed=seed)
        got = _noalloc(n_d, h_d, order, R, L, mode, flags, seed, p_d)
        torch.cuda.synchronize()
        for name, a, b in zip(("hc2ss", "ssidx", "hc2bg", "bgpar"), got, ref):
            if b is None:
                assert a is None
                continue
            assert a.shape == b.shape, name
            assert torch.equal(a, b), f"{name} differs from hd_realize"
        assert bool(torch.isfinite(got[2]).all()) and bool((got[2] > 0).all())


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    lib = _lib.load()
    B, F, R, L = 720000, 40, 500, 10
    assert _lib.size("hd_realize_workspace_size", B, F, R, L, 0) > SCRATCH_BYTES
    one = D.empty(1)
    # refused on the host from the sizes alone: nothing is launched, so the pointers are not read
    rc = lib.hd_realize_noalloc(D.ptr(one), D.ptr(one), D.ptr(one), B, F, R, L, 0, 0, 1e10, 1, None, 0,
                                D.ptr(one), D.ptr(one), D.ptr(one), None, D.stream_ptr())
    assert rc == HD_ERR_SIZE
    assert "scratch" in lib.hd_last_error().decode()
