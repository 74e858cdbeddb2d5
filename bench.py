#!/usr/bin/env python
"""Benchmark: GWB spectra/s of the plain GPU SAM on one MI355X.

A step is one complete spectrum of the full-size workload from its model
parameters: static density with M-Mbulge scatter, Fixed_Time_2PL_SAM
hardening normalisations (Brent), 2PL dynamic binary number, bin integration,
strain, frequency-0 order, 500 Poisson realizations with the 10 loudest
sources per frequency split from the background.  Inputs are parameters and
1-D grids; every per-cell value is computed inside the timed region.

The surrogate half of the metric: the trained MLP (weights shipped in the
package, trained on 20k Latin-hypercube samples of the classic uniform space on
the same grid, 30 frequencies, 500 realizations) predicts log10 h_c median,
5th and 95th percentile spectra for 1e5 uniform parameter vectors, and the
plain GPU SAM recomputes 1% of them as judge.  Reported: surrogate spectra/s
with and without the judge, the judged median error in dex, the plain SAM's
spectra/s at the surrogate's configuration (from the judge runs), and the MFMA
roofline fraction of the inference GEMMs.  An eccentric-SAM line (61x51x61 grid,
40 frequencies, 100 harmonics; deterministic and 200-realization discrete
spectra/s) and a batched-library line (256 of 2048 Latin-hypercube samples at
their own uneven grids, 30 frequencies, 100 realizations: samples/s) follow.

Prints one JSON line (rank 0).  --gpus N>1 is rejected: this build runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GWB spectra/sec on 1 MI355X: plain GPU SAM vs MLP surrogate at ≤0.05 dex err"
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_hd_realize_91x81x101_f40_r500.json")
# wave-level VALU instructions of one k_gz_draw launch (rocprofv3 SQ_INSTS_VALU), for the issue roofline
VALU_FILE = os.path.join(ROOT, "profiles", "pmc_valu_k_gz_draw_91x81x101_f40_r500_l10.json")
# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per SIMD per 4 cycles
VALU_PEAK_GINST_S = 256 * 4 * 2.4 / 4
SURROGATE_WEIGHTS = os.path.join(ROOT, "holodeck_amd", "data", "surrogate_classic_uniform_91x81x101_f30_r500.npz")
BF16_DENSE_PEAK_TFS = 2500.0
CALIB_BYTES = 1 << 30   # copy-kernel calibration size for the PMC passes (bytes read = bytes written)


CPU_PROCS_MAX = 16   # the host-core share of one GPU on the benchmark machines


def cpu_baseline_sample(nz_sample, r_sample):
    """One single-threaded oracle process: a bounded sample of the full-size workload.

    Sample: the same model on an M x Q x nz_sample grid (stage A: density, scatter, Brent norms,
    2PL walk, integration, strain, order) and r_sample realizations of that grid.  Returns the
    timings and the sample's sizes."""
    import oracle
    from oracle.cfg import oracle_cfg
    from oracle.scatter import add_scatter_to_masses
    from holodeck_amd import configs
    full = oracle_cfg(configs.spec_config2())
    M, Q, Z = full['shape']
    cfg = dict(full)
    cfg['shape'] = (M, Q, nz_sample)
    c = cfg['cosmo']
    t0 = time.perf_counter()
    cos = oracle.OracleCosmo(h=c['h'], Om0=c['Om0'])
    mtot, mrat, redz = oracle.edges_from_cfg(cfg)
    dens, zp, gmt = oracle.static_density(mtot, mrat, redz, cfg, cos)
    dens = add_scatter_to_masses(mtot, mrat, dens, cfg['mmbulge']['scatter_dex'])
    dens[zp < 0] = 0.0
    h = cfg['hard']
    mt2, mr2 = np.meshgrid(mtot, mrat, indexing='ij')
    tn0 = time.perf_counter()
    nl = oracle.find_2pwl_norm(h['time'], mt2.ravel(), mr2.ravel(), h['sepa_init'], h['rchar'], h['gamma_inner'],
                               h['gamma_outer'], h['num_steps']).reshape(mt2.shape)
    t_norm = time.perf_counter() - tn0
    fe = cfg['fobs_gw_edges']
    fo = oracle.midpoints(fe) / 2.0
    rf, dn = oracle.dbn_2pwl(fo, h['sepa_init'], h['num_steps'], 10.0 ** nl, h['rchar'], h['gamma_inner'],
                             h['gamma_outer'], dens, mtot, mrat, redz, gmt, cos)
    edges = [mtot, mrat, redz, fe / 2.0]
    num = oracle.integrate_3dx1d(edges, dn)
    h2 = oracle.h2fdf_redz(edges, rf, cos)
    F = num.shape[-1]
    n2, hh = num.reshape(-1, F), h2.reshape(-1, F)
    order = oracle.order_desc(hh[:, 0])
    t_a = time.perf_counter() - t0
    t1 = time.perf_counter()
    oracle.loudest_rng(n2, hh, order, r_sample, 10, 12345)
    t_b = time.perf_counter() - t1
    return dict(t_a=t_a, t_norm=t_norm, t_b=t_b, cells=int(n2.size), M=M, Q=Q, Z=Z, F=F)


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nz_sample=11, r_sample=50, nreals=500, procs=None):
    """The oracle (FP64 C restatement + scipy scatter) as the reference runs library jobs: one
    single-threaded process per host core, each on its own spectrum.  Run before anything touches
    the GPU (the worker processes are started from a process that has not initialised it).

    Each process times the bounded sample (cpu_baseline_sample); its full-size spectrum time is
    extrapolated: stage A by the number of redshift bins (the Brent norms do not depend on z),
    realizations linearly in R (r_sample -> nreals) and in the bins.  Throughput = the sum over
    processes of 1 / (extrapolated spectrum time)."""
    ncpu = os.cpu_count() or 1
    procs = int(procs or min(ncpu, CPU_PROCS_MAX))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", str(nz_sample), str(r_sample)]
    t0 = time.perf_counter()
    ps = [subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
          for _ in range(procs)]
    res = []
    for p in ps:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"cpu baseline worker failed ({p.returncode}): {err[-2000:]}")
        res.append(json.loads(out.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    r0 = res[0]
    M, Q, Z, F = r0['M'], r0['Q'], r0['Z'], r0['F']
    zscale = (Z - 1) / (nz_sample - 1)
    B_full = (M - 1) * (Q - 1) * (Z - 1)
    per = []
    for r in res:
        t_a_full = r['t_norm'] + (r['t_a'] - r['t_norm']) * zscale
        rate = r['cells'] * r_sample / r['t_b']        # cell-realizations per second
        per.append(t_a_full + B_full * F * nreals / rate)
    value = float(sum(1.0 / t for t in per))
    return dict(value=value, unit="spectra/s", cores=procs, kind="port",
                cpu_model=_cpu_model(), host_cpu_count=ncpu,
                sample=(f"oracle (FP64 C + scipy scatter, single-threaded) in {procs} concurrent processes, one per "
                        f"core of {ncpu} ({_cpu_model()}); each ran the full-size model on {M}x{Q}x{nz_sample} edges "
                        f"with R={r_sample}: stage A {np.mean([r['t_a'] for r in res]):.1f}s (scaled x{zscale:.1f} in "
                        f"z, Brent norms unscaled), realizations {np.mean([r['t_b'] for r in res]):.1f}s scaled "
                        f"linearly to R={nreals} over {B_full} bins; {wall:.1f}s wall in all"),
                spectrum_s_mean=float(np.mean(per)), wall_s=wall)


def surrogate_bench(n, judge_frac, seed=7):
    """Surrogate inference on n uniform parameter vectors in the training box, the plain GPU SAM
    judging judge_frac of them (timed pieces and errors)."""
    import torch
    from holodeck_amd import _lib, librarian, surrogate
    net = surrogate.MLPSurrogate.load(SURROGATE_WEIGHTS)
    shape = tuple(int(v) for v in net.meta['shape'])
    F = int(net.meta['fobs_cents'].size)
    ps = librarian.PS_Classic_Phenom_Uniform(sam_shape=shape)
    rng = np.random.default_rng(seed)
    X = net.x_lo + (net.x_hi - net.x_lo) * rng.random((n, net.sizes[0]))
    net.predict(X[:1024])
    torch.cuda.synchronize()
    _lib.timing_enable(True)
    t0 = time.perf_counter()
    net.predict(X)
    t_inf = time.perf_counter() - t0
    gemm_ms, gemm_n = _lib.timing_read("k_mlp_gemm")
    _lib.timing_enable(False)
    flops = 2.0 * n * sum(a * b for a, b in zip(net.sizes[:-1], net.sizes[1:]))
    sw = surrogate.SurrogateWithFallback(net, ps, nfreqs=F, nreals=int(net.meta['nreals']))
    _, rep = sw.predict(X, judge_frac=judge_frac, seed=seed)
    sam_rate = rep["n_judged"] / rep["t_judge_s"] if rep["n_judged"] else None
    tf = flops / (gemm_ms * 1e-3) / 1e12
    return {
        "config": {"workload": (f"MLP {'-'.join(str(v) for v in net.sizes)} (SiLU, BF16 MFMA, FP32 accumulate), "
                                f"trained on {int(net.meta['ntrain'])} classic-uniform samples of the {shape[0]}x"
                                f"{shape[1]}x{shape[2]} grid, {F} frequencies, {int(net.meta['nreals'])} realizations; "
                                f"{n} uniform inference vectors, {judge_frac:.0%} judged by the plain GPU SAM")},
        "spectra_per_s": n / (t_inf + rep["t_judge_s"]),
        "spectra_per_s_inference_only": n / t_inf,
        "plain_sam_spectra_per_s_same_config": sam_rate,
        "x_vs_plain_sam": (n / (t_inf + rep["t_judge_s"])) / sam_rate if sam_rate else None,
        "x_vs_plain_sam_inference_only": (n / t_inf) / sam_rate if sam_rate else None,
        "judged": rep["n_judged"], "judged_median_dex_error": rep["median_dex_error"],
        "judged_p95_dex_error": rep["p95_dex_error"],
        "holdout_median_dex_error": float(net.meta['holdout_median_dex_error']),
        "roofline": {"bound": "mfma", "kernel": "k_mlp_gemm (inference forward)", "achieved": tf,
                     "peak": BF16_DENSE_PEAK_TFS, "unit": "TFLOP/s", "frac": tf / BF16_DENSE_PEAK_TFS,
                     "flops": flops, "gemm_ms": gemm_ms, "launches": gemm_n},
    }


def library_bench(nsamples=256, nreals=100, nfreqs=30, streams=3, nwarm=16):
    """Batched library: nsamples of 2048 Latin-hypercube samples of the classic uniform space, each
    at its own grid drawn from a seed (41-91 x 41-81 x 41-101 edges), 30 frequencies, 100
    realizations, samples pipelined over HIP streams (scatter geometry built ahead by host worker
    processes).  The last nwarm samples run first, untimed (they start the worker pool); the timed
    call covers the first nsamples, whose grids and geometry are all new to it.  Wall clock of the
    timed call: samples/s."""
    from holodeck_amd import librarian as L
    S_all = 2048
    ps = L.PS_Classic_Phenom_Uniform(nsamples=S_all, seed=12345)
    shapes_all = L.sample_shapes(S_all, 777)
    warm = np.arange(S_all - nwarm, S_all)
    L.gwb_library(ps, shapes=shapes_all[warm], nfreqs=nfreqs, nreals=nreals, seed=3, samples=warm, streams=streams)
    rows = np.arange(nsamples)
    shapes = shapes_all[rows]
    t0 = time.perf_counter()
    lib = L.gwb_library(ps, shapes=shapes, nfreqs=nfreqs, nreals=nreals, seed=3, samples=rows, streams=streams)
    dt = time.perf_counter() - t0
    assert np.all(np.isfinite(lib['gwb'])) and not lib['failed'].any()
    cells = float(np.sum(np.prod(shapes - 1, axis=1)))
    return {"config": {"workload": (f"batched library: {nsamples} of 2048 Latin-hypercube samples of the 6-D classic "
                                    "uniform space, each at its own grid (41-91 x 41-81 x 41-101 edges), "
                                    f"{nfreqs} frequencies, {nreals} realizations, {streams} HIP streams; "
                                    f"{nwarm} other samples first, untimed"),
                       "nsamples": nsamples, "nfreqs": nfreqs, "nreals": nreals, "streams": streams},
            "samples_per_s": nsamples / dt, "wall_s": dt,
            "cell_realizations_per_s": cells * nfreqs * nreals / dt}


def eccentric_bench(e0=0.5, nreals=200, nharms=100):
    """Eccentric SAM on the 61x51x61 grid, 40 frequencies i/T, 100 harmonics: the deterministic
    harmonic sum and the discrete variant with nreals Poisson realizations, one spectrum each."""
    import torch
    from holodeck_amd import engine, gravwaves
    from holodeck_amd.sams.sam import Semi_Analytic_Model, evolve_eccen_uniform_single
    from holodeck_amd.constants import PC, YR
    sam = Semi_Analytic_Model(shape=(61, 51, 61))
    fobs = np.arange(1, 41) / (16.03 * YR)
    sepa, ecc = evolve_eccen_uniform_single(sam, e0, 1.0 * PC, 123)
    ins = gravwaves._eccen_inputs(fobs, sam, sepa, ecc)
    engine.gwb_eccen(*ins, nharms)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        det = engine.gwb_eccen(*ins, nharms)
    torch.cuda.synchronize()
    t_det = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    disc = engine.gwb_eccen_discrete(*ins, nharms, nreals, seed=1)
    torch.cuda.synchronize()
    t_disc = time.perf_counter() - t0
    d = det.sum(dim=1)
    mc = disc.sum(dim=1).mean(dim=1)
    return {"config": {"workload": (f"eccentric SAM (default components, KH2013 with 0.3 dex scatter), 61x51x61 grid, "
                                    f"40 frequencies i/T (T=16.03 yr), {nharms} harmonics, e0={e0} at 1 pc, 123 "
                                    f"evolution steps")},
            "deterministic_spectra_per_s": 1.0 / t_det, "discrete_spectra_per_s": 1.0 / t_disc,
            "discrete_nreals": nreals,
            "discrete_mean_over_deterministic": [float(v) for v in (mc / d).cpu().numpy()[:4]]}


def main():
    if len(sys.argv) >= 2 and sys.argv[1] == "--cpu-worker":
        print(json.dumps(cpu_baseline_sample(int(sys.argv[2]), int(sys.argv[3]))), flush=True)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=None,
                    help=f"CPU-baseline processes (default: host cores, at most {CPU_PROCS_MAX})")
    ap.add_argument("--nreals", type=int, default=500)
    ap.add_argument("--streams", type=int, default=2,
                    help="independent spectra in flight on this many HIP streams (consecutive steps alternate)")
    ap.add_argument("--roofline-steps", type=int, default=3,
                    help="single-stream steps after the timed region that time the realization launch")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-eccentric", action="store_true")
    ap.add_argument("--no-library", action="store_true")
    ap.add_argument("--library-n", type=int, default=256)
    ap.add_argument("--surrogate-n", type=int, default=100000)
    ap.add_argument("--judge-frac", type=float, default=0.01)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed steps, copy a 1 GiB device buffer once (reference kernel for PMC byte "
                         "counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(parameter-space samples would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline first: its worker processes start before this process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(procs=args.cpu_procs)

    import torch
    from holodeck_amd import configs, engine, _lib
    from holodeck_amd.cosmology import Cosmology

    spec = configs.spec_config2()
    M, Q, Z = spec['shape']
    F, R, L = spec['nfreq'], args.nreals, 10
    B = (M - 1) * (Q - 1) * (Z - 1)
    fe = configs.fobs_gw_edges(spec)
    cos = Cosmology(h=spec['cosmo']['h'], Om0=spec['cosmo']['Om0'])
    cos.device_tables()

    def step(seed):
        sam = configs.make_sam(spec, cosmo=cos)
        hard = configs.make_hard(spec, sam)
        return sam.gwb_device(fe, hard, R, L, seed=seed)

    # warm-up: the first step on the default stream builds the cached scatter plan and device
    # tables that the streams then only read; the rest alternate over the streams as the timed
    # steps do, so that each stream's memory pool holds its workspaces before the timing starts
    streams = [torch.cuda.Stream() for _ in range(max(1, args.streams))]   # default priority
    for w in range(args.warmup):
        if w == 0:
            out = step(1000 + w)
            torch.cuda.synchronize()
            continue
        with torch.cuda.stream(streams[(w - 1) % len(streams)]):
            out = step(1000 + w)
    torch.cuda.synchronize()
    engine.EVENTS = {}
    _lib.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    outs = []
    for s in range(args.steps):
        with torch.cuda.stream(streams[s % len(streams)]):
            outs.append(step(s))
    torch.cuda.synchronize()
    out = outs[-1]
    dt = time.perf_counter() - t0
    stages = engine.stage_times_ms()
    engine.EVENTS = None
    if len(streams) > 1:
        # roofline pass: with several streams the kernels of different spectra overlap, so the
        # realization launch is timed on its own over a few steps on one stream (after, and
        # outside, the timed steps); enabling clears the records of the timed steps
        _lib.timing_enable(True)
        out_r = None
        for s in range(args.roofline_steps):
            out_r = step(2000 + s)
        torch.cuda.synchronize()
        del out_r
    real_ms, real_n = _lib.timing_read("hd_realize")
    draw_ms, draw_n = _lib.timing_read("k_gz_draw")
    _lib.timing_enable(False)
    hc2bg = out[2]
    assert torch.isfinite(hc2bg).all() and bool((hc2bg > 0).all())
    if args.copy_calibration:
        src = torch.empty(CALIB_BYTES // 8, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ms = dt / args.steps * 1e3
    value = args.steps / dt
    # the realization launch sequence (rank, prep, draw, big-lambda draw, merge), timed by the
    # library with HIP events on its launch stream
    t_real = real_ms / max(real_n, 1) * 1e-3
    alg_bytes = 16.0 * B * F          # number + h2fdf (FP64) read once for all realizations
    achieved = alg_bytes / t_real / 1e9
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as fh:
            traffic = json.load(fh).get("traffic_bytes_per_launch")
    # the draw kernel is bound by instruction issue, not HBM: its VALU instructions per launch
    # (counted by rocprofv3 on the same workload) over its live-timed duration, against the peak
    valu = None
    t_draw = draw_ms / max(draw_n, 1) * 1e-3
    if os.path.exists(VALU_FILE) and t_draw > 0:
        with open(VALU_FILE) as fh:
            vi = json.load(fh).get("valu_insts_per_launch")
        if vi:
            ach = vi / t_draw / 1e9
            valu = {"bound": "valu_issue", "kernel": "k_gz_draw", "achieved": ach, "peak": VALU_PEAK_GINST_S,
                    "unit": "G wave-instructions/s", "frac": ach / VALU_PEAK_GINST_S, "insts_per_launch": vi,
                    "launch_ms": t_draw * 1e3}
    line = {
        "metric": METRIC, "value": value, "unit": "spectra/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": ("full-size plain SAM: 91x81x101 grid (mtot 1e4-1e12 Msol, q 1e-3-1, z 1e-3-10), "
                                "40 frequencies i/T (T=16.03 yr), flat LCDM H0=67.7 Om=0.31, GSMF Schechter + "
                                "GPF/GMT power laws, KH2013 M-Mbulge with 0.28 dex scatter, Fixed_Time_2PL_SAM "
                                "(1 Gyr, -1, +2.5, 100 pc, 1e4 pc), 500 realizations, 10 loudest"),
                   "grid": [M, Q, Z], "nfreq": F, "nreals": R, "nloudest": L,
                   "streams": len(streams)},
        "cell_realizations_per_s": B * F * R * value,
        "stages_ms": stages,
        "roofline": {"bound": "hbm", "kernel": "hd_realize (k_rank_from_order + k_gz_prep + k_gz_draw + k_gz_big + k_realize_group + k_realize_merge)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "alg_bytes_per_launch": alg_bytes, "launch_ms": t_real * 1e3,
                     "launches": real_n, "k_gz_draw_ms": draw_ms / max(draw_n, 1),
                     "timed_on": "timed steps" if len(streams) == 1 else f"{args.roofline_steps} single-stream steps",
                     "valu": valu},
    }
    if not args.no_surrogate and os.path.exists(SURROGATE_WEIGHTS):
        line["surrogate"] = surrogate_bench(args.surrogate_n, args.judge_frac)
    if not args.no_eccentric:
        line["eccentric"] = eccentric_bench()
    if not args.no_library:
        line["library"] = library_bench(args.library_n)
    if cpu is not None:
        line["cpu_baseline"] = cpu
        line["x_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
