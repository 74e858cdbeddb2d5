"""Learned GWB surrogate: an MLP from model parameters to log10 h_c percentile spectra.

Sits beside the reference's Gaussian-process emulator (holodeck/gps/gp_utils.py),
trained on libraries made by :mod:`holodeck_amd.librarian`.  Targets follow the
GP's convention on the library's ``gwb`` array (gp_utils.py:254-349: log10 of
statistics over realizations, with a 1e-40 floor): per frequency the log10 of
the median, 5th and 95th percentile of h_c, i.e. 3 F outputs.

Network: inputs rescaled to [0, 1] over the training box, hidden layers with
SiLU, linear output on standardised targets.  Training and inference run in
the library's HIP kernels (hd_mlp_*): BF16 operands on the matrix cores, FP32
accumulation, FP32 master weights and Adam moments.  PyTorch tensors are the
buffers only.

Outside the training box, or when asked to judge, the plain GPU SAM computes
the spectrum (:class:`SurrogateWithFallback`).
"""
import numpy as np
import torch

from . import _lib
from . import device as D

PERCENTILES = (50.0, 5.0, 95.0)
FLOOR = 1e-40


def targets_from_gwb(gwb):
    """(S, F, R) library h_c -> (S, 3F) log10 [median | 5th | 95th] per frequency."""
    gwb = np.asarray(gwb, dtype=float)
    q = np.percentile(gwb, PERCENTILES, axis=-1)            # (3, S, F)
    return np.log10(np.maximum(q, FLOOR)).transpose(1, 0, 2).reshape(gwb.shape[0], -1)


def _bf16(n_or_shape):
    return torch.empty(n_or_shape, dtype=torch.bfloat16, device=D.require_device())


def _f32(n_or_shape):
    return torch.empty(n_or_shape, dtype=torch.float32, device=D.require_device())


class MLPSurrogate:
    """n_in -> hidden... (SiLU) -> n_out, trained with Adam on mean squared error."""

    def __init__(self, n_in=6, hidden=(256, 256, 256, 256), n_out=90, seed=0):
        D.require_device()
        self.sizes = [int(n_in)] + [int(h) for h in hidden] + [int(n_out)]
        rng = np.random.default_rng(seed)
        self.W, self.b, self.Wb, self.mW, self.vW, self.mb, self.vb = [], [], [], [], [], [], []
        for fan_in, fan_out in zip(self.sizes[:-1], self.sizes[1:]):
            lim = np.sqrt(6.0 / fan_in) if fan_out != self.sizes[-1] else np.sqrt(1.0 / fan_in)
            w = rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)   # (out, in)
            self.W.append(torch.as_tensor(w).to(D.require_device()))
            self.b.append(torch.zeros(fan_out, dtype=torch.float32, device=D.require_device()))
            self.Wb.append(self.W[-1].to(torch.bfloat16))
            self.mW.append(torch.zeros_like(self.W[-1]))
            self.vW.append(torch.zeros_like(self.W[-1]))
            self.mb.append(torch.zeros_like(self.b[-1]))
            self.vb.append(torch.zeros_like(self.b[-1]))
        self.step = 0
        self.x_lo = np.zeros(n_in)
        self.x_hi = np.ones(n_in)
        self.y_mean = np.zeros(n_out)
        self.y_std = np.ones(n_out)

    @property
    def nlayers(self):
        return len(self.sizes) - 1

    # ---- scaling ----
    def _x01(self, X):
        return (np.asarray(X, dtype=float) - self.x_lo) / (self.x_hi - self.x_lo)

    def in_box(self, X):
        x = self._x01(X)
        return np.all((x >= 0.0) & (x <= 1.0), axis=-1)

    # ---- device passes ----
    def _forward(self, xb, keep):
        """xb: (M, n_in) bf16 device.  Returns output f32 (M, n_out) and, if keep, per-layer (z f32, h bf16)."""
        st = D.stream_ptr()
        M = xb.shape[0]
        h = xb
        cache = [(None, xb)]
        for l in range(self.nlayers):
            fin, fout = self.sizes[l], self.sizes[l + 1]
            if l < self.nlayers - 1:
                z = _f32((M, fout)) if keep else None
                hn = _bf16((M, fout))
                _lib.call("hd_mlp_gemm_bf16", D.ptr(h), fin, 1, D.ptr(self.Wb[l]), fin, 1, M, fout, fin,
                          _lib.HD_MLP_EPI_BIAS_SILU, D.ptr(self.b[l]), None, D.ptr(z), D.ptr(hn), fout, st)
                if keep:
                    cache.append((z, hn))
                h = hn
            else:
                y = _f32((M, fout))
                _lib.call("hd_mlp_gemm_bf16", D.ptr(h), fin, 1, D.ptr(self.Wb[l]), fin, 1, M, fout, fin,
                          _lib.HD_MLP_EPI_STORE, D.ptr(self.b[l]), None, D.ptr(y), None, fout, st)
        return y, cache

    def _backward(self, cache, dy):
        """dy: (M, n_out) bf16 gradient of the loss w.r.t. the output.  Returns [(dW f32, db f32)] per layer."""
        st = D.stream_ptr()
        M = dy.shape[0]
        grads = [None] * self.nlayers
        dz = dy
        for l in range(self.nlayers - 1, -1, -1):
            fin, fout = self.sizes[l], self.sizes[l + 1]
            h_in = cache[l][1]                       # (M, fin) bf16
            dW = _f32((fout, fin))
            # dW[o, i] = sum_m dz[m, o] h_in[m, i]
            _lib.call("hd_mlp_gemm_bf16", D.ptr(dz), 1, fout, D.ptr(h_in), 1, fin, fout, fin, M,
                      _lib.HD_MLP_EPI_STORE, None, None, D.ptr(dW), None, fin, st)
            db = _f32(fout)
            _lib.call("hd_mlp_colsum_bf16", D.ptr(dz), M, fout, D.ptr(db), st)
            grads[l] = (dW, db)
            if l > 0:
                # dh[m, i] = sum_o dz[m, o] W[o, i];  dz_prev = dh * silu'(z_prev)
                z_prev = cache[l][0]
                dzp = _bf16((M, fin))
                _lib.call("hd_mlp_gemm_bf16", D.ptr(dz), fout, 1, D.ptr(self.Wb[l]), 1, fin, M, fin, fout,
                          _lib.HD_MLP_EPI_SILU_BWD, None, D.ptr(z_prev), None, D.ptr(dzp), fin, st)
                dz = dzp
        return grads

    def _adam(self, grads, lr, b1=0.9, b2=0.999, eps=1e-8):
        st = D.stream_ptr()
        self.step += 1
        for l, (dW, db) in enumerate(grads):
            _lib.call("hd_mlp_adam", self.W[l].numel(), D.ptr(self.W[l]), D.ptr(dW), D.ptr(self.mW[l]),
                      D.ptr(self.vW[l]), lr, b1, b2, eps, self.step, D.ptr(self.Wb[l]), st)
            _lib.call("hd_mlp_adam", self.b[l].numel(), D.ptr(self.b[l]), D.ptr(db), D.ptr(self.mb[l]),
                      D.ptr(self.vb[l]), lr, b1, b2, eps, self.step, None, st)

    def loss_and_grads(self, X01, Yn):
        """Device loss (MSE on standardised targets) and gradients for one batch; X01, Yn numpy or tensors."""
        xb = D.to_dev(np.asarray(X01, dtype=np.float32), dtype=torch.float32).to(torch.bfloat16).contiguous()
        t = D.to_dev(np.asarray(Yn, dtype=np.float32), dtype=torch.float32)
        y, cache = self._forward(xb, keep=True)
        dy = _bf16(y.shape)
        loss = torch.zeros(1, dtype=torch.float32, device=y.device)
        _lib.call("hd_mlp_mse", D.ptr(y), D.ptr(t), y.numel(), D.ptr(dy), D.ptr(loss), D.stream_ptr())
        return loss, self._backward(cache, dy), y

    # ---- public API ----
    def fit(self, X, Y, epochs=200, batch=1024, lr=1e-3, seed=0, x_lo=None, x_hi=None, verbose=0, lr_min=None,
            rescale=True):
        """Train on parameters X (S, n_in) and targets Y (S, n_out); returns the per-epoch mean loss.

        Adam at ``lr``; with ``lr_min`` the step size follows a cosine from ``lr`` down to ``lr_min``
        over the epochs.  ``rescale=False`` keeps the input box and target scaling of an earlier fit
        (continued training).
        """
        X = np.asarray(X, dtype=float)
        Y = np.asarray(Y, dtype=float)
        if rescale:
            self.x_lo = X.min(axis=0) if x_lo is None else np.asarray(x_lo, float)
            self.x_hi = X.max(axis=0) if x_hi is None else np.asarray(x_hi, float)
            self.y_mean = Y.mean(axis=0)
            self.y_std = np.maximum(Y.std(axis=0), 1e-6)
        X01 = torch.as_tensor(self._x01(X), dtype=torch.float32).to(D.require_device())
        Yn = torch.as_tensor((Y - self.y_mean) / self.y_std, dtype=torch.float32).to(D.require_device())
        S = X.shape[0]
        gen = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            lr_ep = lr if lr_min is None else lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * ep / epochs))
            perm = torch.randperm(S, generator=gen).to(X01.device)
            tot = torch.zeros(1, dtype=torch.float32, device=X01.device)
            nb = 0
            for s0 in range(0, S - batch + 1 if S >= batch else 1, batch):
                idx = perm[s0:s0 + batch]
                xb = X01[idx].to(torch.bfloat16).contiguous()
                t = Yn[idx].contiguous()
                y, cache = self._forward(xb, keep=True)
                dy = _bf16(y.shape)
                _lib.call("hd_mlp_mse", D.ptr(y), D.ptr(t), y.numel(), D.ptr(dy), D.ptr(tot), D.stream_ptr())
                self._adam(self._backward(cache, dy), lr_ep)
                nb += 1
            hist.append(float(tot.item()) / max(nb, 1))
            if verbose and (ep + 1) % verbose == 0:
                print(f"surrogate: epoch {ep + 1}/{epochs} loss {hist[-1]:.4e}", flush=True)
        return np.array(hist)

    def predict_device(self, X01_bf16):
        """Standardised outputs (M, n_out) f32 for device inputs already in [0, 1] (bf16)."""
        y, _ = self._forward(X01_bf16, keep=False)
        return y

    # ---- persistence ----
    def save(self, path, **meta):
        """Weights (FP32 masters), scaling and optional metadata arrays to an ``.npz`` file."""
        arrs = {f"W{l}": self.W[l].cpu().numpy() for l in range(self.nlayers)}
        arrs.update({f"b{l}": self.b[l].cpu().numpy() for l in range(self.nlayers)})
        arrs.update(sizes=np.array(self.sizes), x_lo=self.x_lo, x_hi=self.x_hi, y_mean=self.y_mean,
                    y_std=self.y_std, step=np.array(self.step))
        for k, v in meta.items():
            arrs["meta_" + k] = np.asarray(v)
        np.savez(path, **arrs)

    @classmethod
    def load(cls, path):
        """Model saved by :meth:`save` (no pickled objects); metadata in ``.meta``."""
        with np.load(path, allow_pickle=False) as fh:
            sizes = [int(v) for v in fh["sizes"]]
            net = cls(n_in=sizes[0], hidden=sizes[1:-1], n_out=sizes[-1])
            dev = D.require_device()
            for l in range(net.nlayers):
                net.W[l] = torch.as_tensor(fh[f"W{l}"]).to(dev).contiguous()
                net.b[l] = torch.as_tensor(fh[f"b{l}"]).to(dev).contiguous()
                net.Wb[l] = net.W[l].to(torch.bfloat16)
            net.x_lo, net.x_hi = fh["x_lo"], fh["x_hi"]
            net.y_mean, net.y_std = fh["y_mean"], fh["y_std"]
            net.step = int(fh["step"])
            net.meta = {k[5:]: fh[k] for k in fh.files if k.startswith("meta_")}
        return net

    def predict(self, X, chunk=1 << 17):
        """log10 h_c percentile spectra (M, n_out) for parameters X (M, n_in)."""
        X01 = np.asarray(self._x01(X), dtype=np.float32)
        out = []
        for s0 in range(0, X01.shape[0], chunk):
            xb = D.to_dev(X01[s0:s0 + chunk], dtype=torch.float32).to(torch.bfloat16).contiguous()
            out.append(D.to_host(self.predict_device(xb)).astype(float))
        return np.concatenate(out) * self.y_std + self.y_mean


class SurrogateWithFallback:
    """Surrogate inside the training box; the plain GPU SAM outside it and on a judged fraction.

    The plain SAM answer for parameter rows is the library's: ``gwb_for_params`` over ``pspace``
    with ``nfreqs`` frequencies of a ``pta_dur`` [yr] PTA and ``nreals`` realizations, reduced by
    :func:`targets_from_gwb`.
    """

    def __init__(self, model, pspace, nfreqs=30, nreals=500, pta_dur=None, streams=3):
        from . import librarian
        self.model = model
        self.pspace = pspace
        self.nfreqs = int(nfreqs)
        self.nreals = int(nreals)
        self.pta_dur = librarian.DEF_PTA_DUR if pta_dur is None else pta_dur
        self.streams = streams

    def sam_targets(self, X, seed=0):
        """(n, 3F) log10 percentile spectra of the plain GPU SAM for parameter rows X (n, D)."""
        from . import librarian, utils
        from .constants import YR
        X = np.atleast_2d(np.asarray(X, dtype=float))
        if X.shape[0] == 0:
            return np.zeros((0, self.model.sizes[-1]))
        fe = utils.pta_freqs(dur=self.pta_dur * YR, num=self.nfreqs)[1]
        return librarian.gwb_for_params(self.pspace, X, fobs_edges=fe, nreals=self.nreals, seeds=seed,
                                        streams=self.streams, reduce=targets_from_gwb)

    def predict(self, X, judge_frac=0.01, seed=0):
        """(Y (n, 3F), report).  Report: counts, the judged rows' median / 95th-percentile absolute
        error [dex] over all outputs, and wall-clock seconds of each part."""
        import time
        X = np.asarray(X, dtype=float)
        t0 = time.perf_counter()
        inside = self.model.in_box(X)
        Y = np.empty((X.shape[0], self.model.sizes[-1]))
        if inside.any():
            Y[inside] = self.model.predict(X[inside])
        t1 = time.perf_counter()
        out_idx = np.flatnonzero(~inside)
        if out_idx.size:
            Y[out_idx] = self.sam_targets(X[out_idx], seed=seed)
        t2 = time.perf_counter()
        rng = np.random.default_rng(seed)
        cand = np.flatnonzero(inside)
        nj = int(np.ceil(judge_frac * cand.size)) if cand.size else 0
        judged = np.sort(rng.choice(cand, size=nj, replace=False)) if nj else np.array([], dtype=int)
        err = np.abs(Y[judged] - self.sam_targets(X[judged], seed=seed + 1_000_003)) if nj else None
        t3 = time.perf_counter()
        report = {"n": int(X.shape[0]), "n_fallback": int(out_idx.size), "n_judged": int(nj),
                  "median_dex_error": float(np.median(err)) if nj else None,
                  "p95_dex_error": float(np.percentile(err, 95)) if nj else None,
                  "t_surrogate_s": t1 - t0, "t_fallback_s": t2 - t1, "t_judge_s": t3 - t2, "t_total_s": t3 - t0}
        return Y, report
