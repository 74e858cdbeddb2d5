"""Library generation: parameter distributions and spaces (CPU), and a small batched library on the GPU."""
import numpy as np
import pytest

from holodeck_amd import librarian as L


def test_param_distributions():
    u = L.PD_Uniform("a", -3.5, -1.5)
    assert u(0.0) == -3.5 and u(1.0) == -1.5 and np.isclose(u(0.25), -3.0)
    assert np.allclose(u.extrema, [-3.5, -1.5]) and np.isclose(u.default, -2.5)
    lg = L.PD_Uniform_Log("b", 1e-2, 1e2)
    assert np.isclose(lg(0.5), 1.0) and np.isclose(lg(0.0), 1e-2)
    n = L.PD_Normal("c", 2.0, 0.5, clip=(1.0, 3.0))
    assert np.isclose(n(0.5), 2.0)
    assert np.isclose(n(0.841344746068543), 2.5, rtol=1e-9)   # +1 sigma
    assert n(1e-12) == 1.0 and n(1 - 1e-12) == 3.0              # clipped tails
    assert L.PD_Uniform("d", 0, 1, default=0.3).default == 0.3


def test_classic_uniform_space_is_a_latin_hypercube():
    S = 64
    ps = L.PS_Classic_Phenom_Uniform(nsamples=S, seed=7)
    assert ps.param_samples.shape == (S, 6) and ps.nparameters == 6
    assert ps.param_names == ["gsmf_phi0_log10", "gsmf_mchar0_log10", "mmb_mamp_log10", "mmb_scatter_dex",
                              "hard_time", "hard_gamma_inner"]
    ext = ps.extrema
    assert np.all(ps.param_samples >= ext[:, 0]) and np.all(ps.param_samples <= ext[:, 1])
    # one sample per 1/S stratum in every dimension
    strata = np.floor(ps._uniform_samples * S).astype(int)
    for d in range(6):
        assert np.array_equal(np.sort(strata[:, d]), np.arange(S))
    # seeded: same samples again
    assert np.array_equal(L.PS_Classic_Phenom_Uniform(nsamples=S, seed=7).param_samples, ps.param_samples)
    pd = ps.param_dict(3)
    assert set(pd) == set(ps.param_names) and pd["hard_time"] == ps.param_samples[3, 4]
    with pytest.raises(ValueError, match="not a parameter"):
        ps.model_for_params({"no_such_param": 1.0})
    norm = ps.normalized_params([0.5, None, 0.0, 1.0, 0.5, 0.5])
    assert np.isclose(norm["gsmf_phi0_log10"], -2.5) and np.isclose(norm["gsmf_mchar0_log10"], 11.5)
    assert norm["mmb_mamp_log10"] == 7.6 and norm["mmb_scatter_dex"] == 0.9


def test_sample_shapes():
    sh = L.sample_shapes(500, seed=3)
    assert sh.shape == (500, 3)
    for d, (lo, hi) in enumerate(((41, 91), (41, 81), (41, 101))):
        assert sh[:, d].min() >= lo and sh[:, d].max() <= hi
    assert np.array_equal(sh, L.sample_shapes(500, seed=3))


@pytest.mark.gpu
def test_small_library_on_gpu(gpu, tmp_path):
    """Three samples with uneven grids: finite outputs of the right shapes, written and read back,
    and each sample identical to a direct run_model call with the same seed."""
    print("[library]", flush=True)
    ps = L.PS_Classic_Phenom_Uniform(nsamples=3, seed=11)
    shapes = np.array([[21, 17, 15], [25, 13, 19], [17, 21, 11]])
    path = tmp_path / "lib.npz"
    lib = L.run_library(ps, shapes=shapes, nfreqs=6, nreals=20, nloudest=3, seed=5, path=str(path))
    assert lib["gwb"].shape == (3, 6, 20) and lib["hc_ss"].shape == (3, 6, 20, 3) and lib["hc_bg"].shape == (3, 6, 20)
    for key in ("gwb", "hc_bg", "hc_ss"):
        assert np.all(np.isfinite(lib[key])) and np.all(lib[key] >= 0), key
    assert np.all(lib["gwb"] > 0)
    back = L.load_library(str(path))
    assert np.array_equal(back["hc_bg"], lib["hc_bg"]) and list(back["param_names"].astype(str)) == ps.param_names
    sam, hard = ps.model_for_sample_number(1, sam_shape=tuple(shapes[1]))
    one = L.run_model(sam, hard, nfreqs=6, nreals=20, nloudest=3, seed=5 + 2)
    for key in ("gwb", "hc_ss", "hc_bg"):
        assert np.array_equal(one[key], lib[key][1]), key


@pytest.mark.gpu
def test_pipelined_gwb_library_matches_run_library(gpu):
    """gwb_library over 3 streams, blocks of 3 samples and a 2-entry scatter-plan cache (so plans are
    dropped while other streams may still read them) gives the same GWB, bit for bit, as run_library."""
    print("[gwb_library]", flush=True)
    from holodeck_amd import scatter
    ps = L.PS_Classic_Phenom_Uniform(nsamples=7, seed=13)
    shapes = np.array([[21, 17, 15], [25, 13, 19], [17, 21, 11], [21, 17, 15], [19, 15, 13], [23, 11, 17],
                       [17, 21, 11]])
    old = scatter.MAX_PLANS
    scatter.MAX_PLANS = 2
    try:
        lib = L.gwb_library(ps, shapes=shapes, nfreqs=5, nreals=16, seed=9, streams=3, chunk=3)
        tg = L.gwb_library(ps, shapes=shapes, nfreqs=5, nreals=16, seed=9, streams=2, chunk=4,
                           reduce=lambda hc: np.median(hc, axis=-1), reduce_key='med')
    finally:
        scatter.MAX_PLANS = old
    ref = L.run_library(ps, shapes=shapes, nfreqs=5, nreals=16, seed=9, singles_flag=False)
    assert lib["gwb"].shape == (7, 5, 16) and lib["samples_per_s"] > 0
    assert np.array_equal(lib["gwb"], ref["gwb"])
    assert np.array_equal(tg["med"], np.median(ref["gwb"], axis=-1))


@pytest.mark.gpu
def test_eccentric_library_matches_the_per_model_call(gpu):
    """eccen_library's spectra are the harmonic-summed eccentric GWB of each (sample, e0) model."""
    print("[eccen library]", flush=True)
    from holodeck_amd import gravwaves
    from holodeck_amd.sams.sam import evolve_eccen_uniform_single
    from holodeck_amd.constants import PC, YR
    ps = L.PS_Classic_Phenom_Uniform(nsamples=2, seed=4)
    e0s = [0.0, 0.5, 0.9]
    lib = L.eccen_library(ps, e0s, shape=(13, 9, 11), nfreqs=5, nharms=10)
    assert lib['hc'].shape == (2, 3, 5) and np.all(lib['hc'] > 0)
    sam, _ = ps.model_for_sample_number(1, sam_shape=(13, 9, 11))
    sepa, ecc = evolve_eccen_uniform_single(sam, 0.5, 1.0 * PC, 123)
    g = gravwaves.sam_calc_gwb_single_eccen(lib['fobs'], sam, sepa, ecc, nharms=10)
    assert np.allclose(lib['hc'][1, 1], np.sqrt(g.sum(axis=1)), rtol=1e-12)
    assert np.allclose(lib['fobs'], np.arange(1, 6) / (L.DEF_PTA_DUR * YR))


@pytest.mark.gpu
def test_stream_autotuning_keeps_results(gpu):
    """streams='auto' times each candidate on real rows and keeps the fastest; the results equal a
    fixed-stream run bit for bit."""
    print("[autotune]", flush=True)
    ps = L.PS_Classic_Phenom_Uniform(nsamples=200, sam_shape=(13, 11, 9), seed=2)
    info = {}
    a = L.gwb_for_params(ps, ps.param_samples, fobs_edges=np.arange(1, 6) / 5e8, nreals=8, seeds=3, streams='auto',
                         chunk=64, info=info)
    b = L.gwb_for_params(ps, ps.param_samples, fobs_edges=np.arange(1, 6) / 5e8, nreals=8, seeds=3, streams=1)
    assert info["streams"] in L.AUTOTUNE_STREAMS and len(info["probe_rates"]) == len(L.AUTOTUNE_STREAMS)
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_run_library_isolates_failures_and_resumes(gpu, tmp_path, monkeypatch):
    """A sample whose run raises is marked failed (NaN outputs and parameters, a 'fail' file), the
    others are unaffected; a rerun reads finished samples back and recomputes only failed or missing
    ones (gen_lib.py:283-320, combine.py:206-212), with the same results."""
    print("[library failures and resume]", flush=True)
    ps = L.PS_Classic_Phenom_Uniform(nsamples=4, seed=21)
    shapes = np.array([[15, 13, 11], [13, 11, 9], [17, 13, 11], [13, 13, 9]])
    kw = dict(shapes=shapes, nfreqs=5, nreals=12, nloudest=2, seed=4)
    ref = L.run_library(ps, **kw)
    assert not ref["failed"].any()
    real = L.run_model
    calls = []

    def flaky(sam, hard, **k):
        calls.append(k["seed"])
        if k["seed"] == 4 + 2 * 2:            # sample 2
            raise RuntimeError("planted failure")
        return real(sam, hard, **k)

    monkeypatch.setattr(L, "run_model", flaky)
    sims = tmp_path / "sims"
    lib = L.run_library(ps, sims_dir=str(sims), **kw)
    assert list(lib["failed"]) == [False, False, True, False]
    assert "planted failure" in lib["fail_messages"][2]
    assert np.all(np.isnan(lib["gwb"][2])) and np.all(np.isnan(lib["sample_params"][2]))
    for i in (0, 1, 3):
        for key in ("gwb", "hc_ss", "hc_bg"):
            assert np.array_equal(lib[key][i], ref[key][i]), (i, key)
    with np.load(L.sim_fname(str(sims), 2)) as fh:
        assert fh.files == ["fail"]
    # resume: sample 2 (failed) and sample 1 (file removed) are recomputed, 0 and 3 read back
    (sims / "sim_000001.npz").unlink()
    monkeypatch.setattr(L, "run_model", lambda sam, hard, **k: (calls.append(k["seed"]), real(sam, hard, **k))[1])
    calls.clear()
    again = L.run_library(ps, sims_dir=str(sims), **kw)
    assert sorted(calls) == [4 + 2 * 1, 4 + 2 * 2] and int(again["resumed"]) == 2
    assert not again["failed"].any()
    for key in ("gwb", "hc_ss", "hc_bg"):
        assert np.array_equal(again[key], ref[key]), key


@pytest.mark.gpu
def test_gwb_library_isolates_failures_and_resumes(gpu, tmp_path, monkeypatch):
    """The pipelined library: a row whose model raises fails alone (NaN, flagged); the checkpoint
    written after every block lets an interrupted run continue with identical results."""
    print("[gwb_library failures and resume]", flush=True)
    ps = L.PS_Classic_Phenom_Uniform(nsamples=9, seed=23)
    shapes = np.array([[13, 11, 9], [15, 11, 9], [13, 13, 11], [11, 11, 9], [13, 11, 11], [15, 13, 9],
                       [11, 13, 9], [13, 11, 9], [15, 11, 11]])
    kw = dict(shapes=shapes, nfreqs=4, nreals=10, seed=6, streams=2, chunk=3)
    ref = L.gwb_library(ps, **kw)
    assert not ref["failed"].any() and np.all(np.isfinite(ref["gwb"]))
    real_hard = ps._init_hard
    bad = ps.param_samples[4]

    def flaky_hard(sam, settings):
        if all(settings[n] == v for n, v in zip(ps.param_names, bad)):
            raise ValueError("planted failure")
        return real_hard(sam, settings)

    monkeypatch.setattr(ps, "_init_hard", flaky_hard)
    lib = L.gwb_library(ps, **kw)
    assert list(np.flatnonzero(lib["failed"])) == [4] and "planted failure" in lib["fail_messages"][4]
    assert np.all(np.isnan(lib["gwb"][4])) and np.all(np.isnan(lib["sample_params"][4]))
    ok = np.arange(9) != 4
    assert np.array_equal(lib["gwb"][ok], ref["gwb"][ok])
    # an interrupted run: the second block raises (outside any row), leaving the first block's checkpoint
    monkeypatch.setattr(ps, "_init_hard", real_hard)
    path = str(tmp_path / "lib.npz")
    real_on = L.gwb_for_params

    def interrupted(*a, **k):
        on_block = k["on_block"]
        seen = []

        def once(start, hc, fl, ms):
            on_block(start, hc, fl, ms)
            seen.append(start)
            if len(seen) == 1:
                raise KeyboardInterrupt
        k["on_block"] = once
        return real_on(*a, **k)

    monkeypatch.setattr(L, "gwb_for_params", interrupted)
    with pytest.raises(KeyboardInterrupt):
        L.gwb_library(ps, path=path, **kw)
    with np.load(path, allow_pickle=False) as fh:
        assert fh["done"].sum() == 3
    monkeypatch.setattr(L, "gwb_for_params", real_on)
    done = L.gwb_library(ps, path=path, **kw)
    assert done["done"].all() and np.array_equal(done["gwb"], ref["gwb"])
    with pytest.raises(ValueError, match="another library"):
        L.gwb_library(ps, path=path, **dict(kw, seed=7))
