"""Surrogate workflow from the command line: make a training library, train, evaluate.

    python -m holodeck_amd.surrogate_cli generate --nsamples 20000 --part 0 --parts 4 --out lib0.npz
    python -m holodeck_amd.surrogate_cli train --library lib0.npz lib1.npz ... --out weights.npz
    python -m holodeck_amd.surrogate_cli evaluate --weights weights.npz --ninfer 100000

    python -m holodeck_amd.surrogate_cli generate-eccen --nsamples 256 --ne0 20 --out ecc.npz

``generate`` runs Latin-hypercube samples of the classic uniform phenomenological space
(librarian/param_spaces_classic.py:100-117) on one grid through the pipelined GPU SAM and
stores parameters and log10 percentile targets (median, 5th, 95th h_c per frequency).  A
library can be made in parts (``--part i --parts k`` takes every k-th sample from i).
``generate-eccen`` runs samples x initial eccentricities through the eccentric SAM (61x51x61,
40 frequencies, 100 harmonics, deterministic harmonic sums) for a 7-input surrogate of log10 h_c.
``train`` holds out a seeded random subset (whole samples with ``--group``), fits the MLP with
the HIP kernels and reports the held-out error in dex.  ``evaluate`` times inference on uniform random parameter vectors,
with the plain GPU SAM judging a fraction of them.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import librarian, surrogate


def _space(nsamples, seed, shape):
    return librarian.PS_Classic_Phenom_Uniform(nsamples=nsamples, sam_shape=tuple(shape), seed=seed)


def cmd_generate(a):
    ps = _space(a.nsamples, a.seed, a.shape)
    samples = np.arange(a.part, a.nsamples, a.parts)
    lib = librarian.gwb_library(ps, nfreqs=a.nfreqs, nreals=a.nreals, seed=a.seed, samples=samples,
                                streams=a.streams, reduce=surrogate.targets_from_gwb, log_every=a.log_every,
                                path=a.out + ".partial.npz")   # checkpoint: a rerun resumes
    keep = ~lib['failed']   # failed samples (NaN outputs) are left out of the training set
    np.savez(a.out, samples=samples[keep], sample_params=lib['sample_params'][keep],
             targets=lib['targets'][keep].astype(np.float32), param_names=lib['param_names'],
             fobs_cents=lib['fobs_cents'], nreals=np.array(a.nreals), shape=np.array(a.shape),
             extrema=ps.extrema, lhs_seed=np.array(a.seed), lhs_nsamples=np.array(a.nsamples))
    print(json.dumps({"generated": int(keep.sum()), "failed": int((~keep).sum()), "samples_per_s": lib['samples_per_s'],
                      "grid": list(a.shape), "nfreqs": a.nfreqs, "nreals": a.nreals}), flush=True)


def cmd_generate_eccen(a):
    """Eccentric library for a 7-input surrogate: samples x initial eccentricities, targets log10
    h_c at each frequency (deterministic harmonic sums)."""
    ps = _space(a.nsamples, a.seed, a.shape)
    e0 = np.linspace(0.0, 0.99, a.ne0)
    lib = librarian.eccen_library(ps, e0, shape=a.shape, nfreqs=a.nfreqs, nharms=a.nharms, log_every=a.log_every)
    S, E, F = lib['hc'].shape
    X = np.concatenate([np.repeat(lib['sample_params'], E, axis=0), np.tile(e0, S)[:, None]], axis=1)
    Y = np.log10(np.maximum(lib['hc'].reshape(S * E, F), surrogate.FLOOR))
    extrema = np.vstack([ps.extrema, [[0.0, 0.99]]])
    np.savez(a.out, sample_params=X, targets=Y.astype(np.float32),
             param_names=np.array(list(ps.param_names) + ['eccen_init']).astype('S'), fobs_cents=lib['fobs'],
             nreals=np.array(0), shape=np.array(a.shape), extrema=extrema, nharms=np.array(a.nharms))
    print(json.dumps({"generated": int(S * E), "samples": int(S), "eccen_inits": int(E),
                      "samples_per_s": lib['samples_per_s'], "spectra_per_s": lib['samples_per_s'] * E}), flush=True)


def _load(paths):
    X, Y, meta = [], [], None
    for p in paths:
        with np.load(p, allow_pickle=False) as fh:
            X.append(fh['sample_params'])
            Y.append(fh['targets'].astype(float))
            if meta is None:
                meta = {k: fh[k] for k in ('param_names', 'fobs_cents', 'nreals', 'shape', 'extrema')}
    return np.concatenate(X), np.concatenate(Y), meta


def cmd_train(a):
    X, Y, meta = _load(a.library)
    rng = np.random.default_rng(a.seed)
    # hold out whole groups of consecutive rows (a parameter sample's rows when a library has
    # several per sample, e.g. one per initial eccentricity)
    g = max(1, int(a.group))
    groups = rng.permutation(X.shape[0] // g)
    nh = max(1, a.holdout // g)
    rows = np.arange(X.shape[0]).reshape(-1, g)
    hold, train = rows[groups[:nh]].ravel(), rows[groups[nh:]].ravel()
    net = surrogate.MLPSurrogate(n_in=X.shape[1], hidden=[a.width] * a.depth, n_out=Y.shape[1], seed=a.seed)
    ext = meta['extrema']
    t0 = time.perf_counter()
    hist = net.fit(X[train], Y[train], epochs=a.epochs, batch=a.batch, lr=a.lr, lr_min=a.lr_min, seed=a.seed,
                   x_lo=ext[:, 0], x_hi=ext[:, 1], verbose=a.log_every)
    t_fit = time.perf_counter() - t0
    err = np.abs(net.predict(X[hold]) - Y[hold])
    rep = {"train": int(train.size), "holdout": int(hold.size), "epochs": a.epochs, "fit_s": t_fit,
           "final_loss": float(hist[-1]), "median_dex_error": float(np.median(err)),
           "p95_dex_error": float(np.percentile(err, 95))}
    if int(meta['nreals']) > 0:                 # percentile targets: [median | p05 | p95] per frequency
        F = Y.shape[1] // 3
        rep["median_dex_error_by_output"] = {name: float(np.median(err[:, k * F:(k + 1) * F]))
                                             for k, name in enumerate(("p50", "p05", "p95"))}
    net.save(a.out, param_names=meta['param_names'], fobs_cents=meta['fobs_cents'], nreals=meta['nreals'],
             shape=meta['shape'], holdout_median_dex_error=rep["median_dex_error"],
             holdout_p95_dex_error=rep["p95_dex_error"], ntrain=train.size)
    print(json.dumps(rep), flush=True)


def cmd_evaluate(a):
    net = surrogate.MLPSurrogate.load(a.weights)
    shape = tuple(int(v) for v in net.meta['shape'])
    ps = _space(None, None, shape)
    F = int(net.meta['fobs_cents'].size)
    rng = np.random.default_rng(a.seed)
    X = net.x_lo + (net.x_hi - net.x_lo) * rng.random((a.ninfer, net.sizes[0]))
    net.predict(X[:1024])                      # first-call set-up outside the timing
    sw = surrogate.SurrogateWithFallback(net, ps, nfreqs=F, nreals=int(net.meta['nreals']))
    Y, rep = sw.predict(X, judge_frac=a.judge_frac, seed=a.seed)
    rep["spectra_per_s_with_judge"] = a.ninfer / rep["t_total_s"]
    rep["spectra_per_s_surrogate_only"] = a.ninfer / rep["t_surrogate_s"]
    print(json.dumps(rep), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m holodeck_amd.surrogate_cli")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("generate")
    g.add_argument("--nsamples", type=int, default=20000)
    g.add_argument("--seed", type=int, default=20240)
    g.add_argument("--part", type=int, default=0)
    g.add_argument("--parts", type=int, default=1)
    g.add_argument("--shape", type=int, nargs=3, default=(91, 81, 101))
    g.add_argument("--nfreqs", type=int, default=30)
    g.add_argument("--nreals", type=int, default=500)
    g.add_argument("--streams", type=int, default=3)
    g.add_argument("--log-every", type=int, default=256)
    g.add_argument("--out", required=True)
    ge = sub.add_parser("generate-eccen")
    ge.add_argument("--nsamples", type=int, default=256)
    ge.add_argument("--seed", type=int, default=5150)
    ge.add_argument("--ne0", type=int, default=20)
    ge.add_argument("--shape", type=int, nargs=3, default=(61, 51, 61))
    ge.add_argument("--nfreqs", type=int, default=40)
    ge.add_argument("--nharms", type=int, default=100)
    ge.add_argument("--log-every", type=int, default=32)
    ge.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--library", nargs="+", required=True)
    t.add_argument("--holdout", type=int, default=1000)
    t.add_argument("--group", type=int, default=1, help="rows per parameter sample (held out together)")
    t.add_argument("--width", type=int, default=256)
    t.add_argument("--depth", type=int, default=4)
    t.add_argument("--epochs", type=int, default=2000)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-min", type=float, default=1e-5)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--log-every", type=int, default=100)
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", required=True)
    e.add_argument("--ninfer", type=int, default=100000)
    e.add_argument("--judge-frac", type=float, default=0.01)
    e.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    {"generate": cmd_generate, "generate-eccen": cmd_generate_eccen, "train": cmd_train,
     "evaluate": cmd_evaluate}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
