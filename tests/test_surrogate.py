"""Learned surrogate: targets (CPU); BF16 MFMA GEMM, gradients and training (GPU) against float32 references."""
import numpy as np
import pytest

from holodeck_amd import surrogate as S


def test_targets_from_gwb():
    rng = np.random.default_rng(0)
    g = 10.0 ** rng.uniform(-16, -14, size=(3, 5, 101))
    g[1, 2] = 0.0
    t = S.targets_from_gwb(g)
    assert t.shape == (3, 15)
    assert np.allclose(t[0, :5], np.log10(np.median(g[0], axis=-1)))
    assert np.allclose(t[2, 5:10], np.log10(np.percentile(g[2], 5, axis=-1)))
    assert np.allclose(t[0, 10:], np.log10(np.percentile(g[0], 95, axis=-1)))
    assert t[1, 2] == -40.0     # floor 1e-40, as the GP targets


def _bf16_round(a):
    import torch
    return torch.as_tensor(np.asarray(a, dtype=np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("M,N,K,trans", [(64, 64, 32, "nt"), (200, 90, 256, "nt"), (1000, 256, 6, "nt"),
                                         (256, 6, 1000, "tn"), (130, 77, 45, "nn")])
def test_mlp_gemm_matches_fp32(M, N, K, trans, gpu):
    """C = A B^T on the matrix cores vs float32 numpy on the same bf16-rounded operands, for the three stride
    patterns the network uses (forward, weight gradient, input gradient)."""
    import torch
    from holodeck_amd import _lib
    from holodeck_amd import device as D
    print(f"[mlp gemm {M}x{N}x{K} {trans}]", flush=True)
    rng = np.random.default_rng(M + N + K)
    A = _bf16_round(rng.normal(size=(M, K)))
    B = _bf16_round(rng.normal(size=(N, K)))
    ref = A.astype(np.float64) @ B.astype(np.float64).T
    if trans == "nt":      # A (M,K) row-major, B (N,K) row-major
        a_store, sam, sak = A, K, 1
        b_store, sbn, sbk = B, K, 1
    elif trans == "tn":    # A stored (K,M), B stored (K,N)
        a_store, sam, sak = np.ascontiguousarray(A.T), 1, M
        b_store, sbn, sbk = np.ascontiguousarray(B.T), 1, N
    else:                  # A row-major, B stored (K,N)
        a_store, sam, sak = A, K, 1
        b_store, sbn, sbk = np.ascontiguousarray(B.T), 1, N
    ad = D.to_dev(a_store, dtype=torch.float32).to(torch.bfloat16).contiguous()
    bd = D.to_dev(b_store, dtype=torch.float32).to(torch.bfloat16).contiguous()
    c = torch.zeros((M, N), dtype=torch.float32, device=ad.device)
    _lib.call("hd_mlp_gemm_bf16", D.ptr(ad), sam, sak, D.ptr(bd), sbn, sbk, M, N, K, _lib.HD_MLP_EPI_STORE, None,
              None, D.ptr(c), None, N, D.stream_ptr())
    got = D.to_host(c)
    # products of bf16 values are exact in fp32; only the fp32 summation order differs
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5 * np.sqrt(K) * np.abs(ref).max()), np.abs(got - ref).max()


@pytest.mark.gpu
def test_mlp_forward_backward_match_fp32_autograd(gpu):
    """Loss and every weight/bias gradient of one batch vs torch float32 autograd on the CPU with the same
    bf16 weights and inputs (tolerance from the bf16 rounding of activations and back-propagated errors)."""
    import torch
    print("[mlp gradients]", flush=True)
    rng = np.random.default_rng(3)
    net = S.MLPSurrogate(n_in=6, hidden=(64, 64), n_out=12, seed=1)
    X = rng.uniform(size=(256, 6))
    Y = rng.normal(size=(256, 12))
    loss, grads, y = net.loss_and_grads(X, Y)
    # reference
    Ws = [torch.tensor(w.to(torch.bfloat16).float().cpu().numpy(), requires_grad=True) for w in net.W]
    bs = [torch.tensor(b.cpu().numpy(), requires_grad=True) for b in net.b]
    h = torch.tensor(_bf16_round(X))
    for l in range(net.nlayers):
        z = h @ Ws[l].T + bs[l]
        h = torch.nn.functional.silu(z) if l < net.nlayers - 1 else z
    ref_loss = torch.mean((h - torch.tensor(Y, dtype=torch.float32)) ** 2)
    ref_loss.backward()
    assert np.allclose(y.cpu().numpy(), h.detach().numpy(), rtol=3e-2, atol=3e-2)
    assert abs(float(loss.item()) - float(ref_loss)) < 2e-2 * float(ref_loss)
    for l, (dW, db) in enumerate(grads):
        gW, gb = Ws[l].grad.numpy(), bs[l].grad.numpy()
        for got, ref, what in ((dW.cpu().numpy(), gW, "W"), (db.cpu().numpy(), gb, "b")):
            scale = np.abs(ref).max()
            err = np.abs(got - ref).max() / scale
            cos = np.dot(got.ravel(), ref.ravel()) / (np.linalg.norm(got) * np.linalg.norm(ref))
            assert err < 0.05 and cos > 0.999, (l, what, err, cos)


@pytest.mark.gpu
def test_mlp_training_learns_a_smooth_map(gpu):
    """Adam on the device fits a smooth 6 -> 8 map to a few percent of its spread, and the trained network
    beats predicting the training mean by a wide margin on held-out points."""
    print("[mlp training]", flush=True)
    rng = np.random.default_rng(5)
    def f(X):
        k = np.arange(1, 9)
        return np.sin(X @ np.linspace(0.5, 2.0, 6))[:, None] * np.log(k)[None, :] + (X[:, :1] ** 2) * k[None, :] / 8
    Xtr, Xte = rng.uniform(size=(4096, 6)), rng.uniform(size=(1024, 6))
    net = S.MLPSurrogate(n_in=6, hidden=(128, 128, 128), n_out=8, seed=2)
    hist = net.fit(Xtr, f(Xtr), epochs=120, batch=256, lr=2e-3, seed=0, x_lo=np.zeros(6), x_hi=np.ones(6))
    assert hist[-1] < 0.05 * hist[0], hist[[0, -1]]
    pred = net.predict(Xte)
    truth = f(Xte)
    rms = np.sqrt(np.mean((pred - truth) ** 2))
    rms0 = np.sqrt(np.mean((truth - f(Xtr).mean(axis=0)) ** 2))
    assert rms < 0.1 * rms0, (rms, rms0)
    assert np.all(net.in_box(Xte)) and not net.in_box(np.full((1, 6), 1.5))[0]


@pytest.mark.gpu
def test_surrogate_on_a_sam_library_gains_over_the_plain_sam(gpu, tmp_path):
    """What the trained network gains: on a small SAM library (17x15x13 grid, 6 frequencies, 64
    realizations) the MLP's held-out error is a fraction of the no-model guess (the training mean),
    it answers far faster than the plain GPU SAM, it survives save/load bit for bit, and points
    outside the training box fall back to the plain SAM exactly."""
    import time
    from holodeck_amd import librarian as L
    print("[surrogate on SAM library]", flush=True)
    ps = L.PS_Classic_Phenom_Uniform(nsamples=640, sam_shape=(17, 15, 13), seed=21)
    t0 = time.perf_counter()
    lib = L.gwb_library(ps, nfreqs=6, nreals=64, seed=3, reduce=S.targets_from_gwb)
    t_sam = (time.perf_counter() - t0) / 640
    X, Y = lib["sample_params"], lib["targets"]
    tr, te = np.arange(576), np.arange(576, 640)
    net = S.MLPSurrogate(n_in=6, hidden=(128, 128, 128), n_out=18, seed=1)
    ext = ps.extrema
    net.fit(X[tr], Y[tr], epochs=400, batch=64, lr=2e-3, lr_min=1e-4, seed=0, x_lo=ext[:, 0], x_hi=ext[:, 1])
    err = np.median(np.abs(net.predict(X[te]) - Y[te]))
    err0 = np.median(np.abs(Y[tr].mean(axis=0) - Y[te]))
    print(f"held-out median error {err:.4f} dex, mean-guess {err0:.4f} dex", flush=True)
    assert err < 0.25 * err0, (err, err0)
    Xq = np.repeat(X[te], 100, axis=0)
    net.predict(Xq[:64])
    t0 = time.perf_counter()
    net.predict(Xq)
    t_mlp = (time.perf_counter() - t0) / Xq.shape[0]
    assert t_sam > 20 * t_mlp, (t_sam, t_mlp)
    path = tmp_path / "w.npz"
    net.save(str(path), note="x")
    back = S.MLPSurrogate.load(str(path))
    assert np.array_equal(back.predict(X[te]), net.predict(X[te])) and str(back.meta["note"]) == "x"
    sw = S.SurrogateWithFallback(net, ps, nfreqs=6, nreals=64)
    Xo = np.vstack([X[te[:4]], ext[:, 1] + 0.1])          # last row outside the box
    Yo, rep = sw.predict(Xo, judge_frac=0.5, seed=11)
    assert rep["n_fallback"] == 1 and rep["n_judged"] == 2 and rep["median_dex_error"] is not None
    assert np.array_equal(Yo[-1], sw.sam_targets(Xo[-1:], seed=11)[0])
    assert np.array_equal(Yo[:4], net.predict(Xo[:4]))
