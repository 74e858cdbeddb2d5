"""Library generation from the command line (librarian/gen_lib.py:50-337, one GPU).

    python -m holodeck_amd.gen_lib --nsamples 2048 --nfreqs 30 --nreals 100 --uneven --out lib.npz

Latin-hypercube samples of the classic uniform phenomenological space; with ``--uneven`` each
sample gets its own grid (edge counts uniform in mtot 41-91, mrat 41-81, redz 41-101 from
``--shape-seed``).  Only the independent Poisson GWB is made (``gwb``, (S, F, R)), on the
pipelined device path; ``--full`` runs ``run_model`` per sample instead (loud sources and
background too).  A sample that fails is marked (``failed``, ``fail_messages``, NaN outputs and
parameters) and the rest go on; ``--out`` is written after every block and a rerun resumes it
(``--recreate`` starts over; with ``--full``, ``--sims-dir`` keeps one file per sample).  The reference's MPI scatter of sample numbers over processes becomes one
process per GPU; this build runs on one GPU.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import librarian


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m holodeck_amd.gen_lib")
    ap.add_argument("--nsamples", type=int, default=2048)
    ap.add_argument("--seed", type=int, default=12345, help="Latin-hypercube and Philox seed")
    ap.add_argument("--shape", type=int, nargs=3, default=(91, 81, 101), help="grid edges without --uneven")
    ap.add_argument("--uneven", action="store_true")
    ap.add_argument("--shape-seed", type=int, default=777)
    ap.add_argument("--nfreqs", type=int, default=30)
    ap.add_argument("--nreals", type=int, default=100)
    ap.add_argument("--nloudest", type=int, default=5)
    ap.add_argument("--pta-dur", type=float, default=librarian.DEF_PTA_DUR, help="[yr]")
    ap.add_argument("--streams", default="auto", help="HIP streams, or 'auto' (timed on the first samples)")
    ap.add_argument("--full", action="store_true", help="run_model per sample (gwb, hc_ss, hc_bg)")
    ap.add_argument("--log-every", type=int, default=256)
    ap.add_argument("--out", default=None,
                    help="write the library (.npz); it is also the checkpoint a rerun resumes from")
    ap.add_argument("--recreate", action="store_true", help="recompute everything instead of resuming")
    ap.add_argument("--sims-dir", default=None,
                    help="with --full: one file per sample there (resumed as gen_lib.py:283-294)")
    a = ap.parse_args(argv)
    ps = librarian.PS_Classic_Phenom_Uniform(nsamples=a.nsamples, sam_shape=tuple(a.shape), seed=a.seed)
    shapes = librarian.sample_shapes(a.nsamples, a.shape_seed) if a.uneven else None
    t0 = time.perf_counter()
    if a.full:
        lib = librarian.run_library(ps, shapes=shapes, pta_dur=a.pta_dur, nfreqs=a.nfreqs, nreals=a.nreals,
                                    nloudest=a.nloudest, seed=a.seed, log_every=a.log_every, sims_dir=a.sims_dir,
                                    recreate=a.recreate)
        rate = a.nsamples / (time.perf_counter() - t0)
        used = 1
        computed = a.nsamples - int(lib['resumed'])
    else:
        streams = a.streams if a.streams == "auto" else int(a.streams)
        lib = librarian.gwb_library(ps, shapes=shapes, pta_dur=a.pta_dur, nfreqs=a.nfreqs, nreals=a.nreals,
                                    seed=a.seed, streams=streams, log_every=a.log_every, path=a.out,
                                    resume=not a.recreate)
        rate = lib.pop('samples_per_s')
        used = int(lib.pop('streams_used'))
        computed = int(a.nsamples)
    if shapes is not None:
        lib['sam_shapes'] = shapes
    cells = (np.prod(shapes - 1, axis=1).sum() if shapes is not None
             else a.nsamples * np.prod(np.array(a.shape) - 1))
    gwb = lib['gwb']
    good = ~lib['failed']
    ok = bool(np.all(np.isfinite(gwb[good])) and np.all(gwb[good] >= 0))
    if a.out:
        np.savez(a.out, **lib)
    print(json.dumps({"samples": a.nsamples, "computed": computed, "failed": int((~good).sum()),
                      "samples_per_s": rate, "uneven": bool(a.uneven),
                      "nfreqs": a.nfreqs, "nreals": a.nreals, "full": bool(a.full), "finite": ok, "streams": used,
                      "cell_realizations_per_s": float(cells) * a.nfreqs * a.nreals * rate / a.nsamples,
                      "median_hc_f0": float(np.nanmedian(gwb[:, 0, :]))}), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
