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
