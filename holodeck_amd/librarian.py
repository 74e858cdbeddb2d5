"""Library generation over a parameter space, on the device path.

Mirrors holodeck.librarian: the parameter distributions and the Latin-hypercube
parameter space (librarian/lib_tools.py:34-465, 467-600), the classic
phenomenological space (librarian/param_spaces_classic.py:9-108), ``run_model``
(lib_tools.py:714-842) and the combined library file (librarian/combine.py:
86-270; written as ``.npz`` with the same dataset and attribute names, since
h5py is not a dependency here).

Each sample is one Semi_Analytic_Model + Fixed_Time_2PL_SAM run on the GPU:
density (with M-Mbulge scatter) -> hardening norms -> evolution -> number ->
strain -> loud-source split (``hc_ss``, ``hc_bg``) and, as the reference's
``gwb_flag``, an independent Poisson GWB (``gwb``).  Samples may have different
grid shapes.  ``run_library`` runs them back to back with every output of
``run_model``; ``gwb_library`` keeps only the GWB and overlaps samples on
several HIP streams, with the scatter geometry of new grids built ahead on host
threads.
"""
import abc
import collections
import time

import numpy as np
import scipy.stats

from .constants import GYR, PC, YR
from . import host_relations, hardening, utils
from . import sams
from .sams.sam import Semi_Analytic_Model

DEF_PTA_DUR = 16.03        # [yr]  (librarian/__init__.py)
DEF_NUM_FBINS = 40
DEF_NUM_REALS = 100
DEF_NUM_LOUDEST = 5


# ---------------------------------------------------------------------------
# parameter distributions (lib_tools.py:467-600)
# ---------------------------------------------------------------------------
class _Param_Dist(abc.ABC):
    """Maps uniform [0, 1] values to parameter values (usually an inverse CDF)."""

    def __init__(self, name, default=None, clip=None):
        if clip is not None:
            assert len(clip) == 2
        self._clip = clip
        self._name = name
        self._default = default

    def __call__(self, xx):
        rv = self._dist_func(xx)
        if self._clip is not None:
            rv = np.clip(rv, *self._clip)
        return rv

    @abc.abstractmethod
    def _dist_func(self, xx):
        pass

    @property
    def extrema(self):
        return self(np.asarray([0.0, 1.0]))

    @property
    def name(self):
        return self._name

    @property
    def default(self):
        """The fixed default if one was given, else the value at uniform 0.5."""
        if self._default is not None:
            return self._default
        return self(0.5)


class PD_Uniform(_Param_Dist):
    def __init__(self, name, lo, hi, **kwargs):
        super().__init__(name, **kwargs)
        self._lo = lo
        self._hi = hi

    def _dist_func(self, xx):
        return self._lo + (self._hi - self._lo) * np.asarray(xx)


class PD_Uniform_Log(_Param_Dist):
    def __init__(self, name, lo, hi, **kwargs):
        super().__init__(name, **kwargs)
        assert lo > 0.0 and hi > 0.0
        self._lo_log10 = np.log10(lo)
        self._hi_log10 = np.log10(hi)

    def _dist_func(self, xx):
        return np.power(10.0, self._lo_log10 + (self._hi_log10 - self._lo_log10) * np.asarray(xx))


class PD_Normal(_Param_Dist):
    """Normal distribution with the given mean and standard deviation (use ``clip`` against extremes)."""

    def __init__(self, name, mean, stdev, clip=None, **kwargs):
        assert stdev > 0.0
        super().__init__(name, clip=clip, **kwargs)
        self._mean = mean
        self._stdev = stdev
        self._frozen = scipy.stats.norm(loc=mean, scale=stdev)

    def _dist_func(self, xx):
        return self._frozen.ppf(xx)


# ---------------------------------------------------------------------------
# parameter spaces (lib_tools.py:34-465, param_spaces_classic.py)
# ---------------------------------------------------------------------------
class _Param_Space(abc.ABC):
    """A parameter space: distributions per varied parameter, Latin-hypercube samples, and the
    SAM/hardening builders for a parameter dictionary (unvaried parameters from ``DEFAULTS``)."""

    DEFAULTS = {}

    def __init__(self, parameters, nsamples=None, sam_shape=None, seed=None):
        parameters = list(parameters)
        if len(parameters) == 0 or not all(isinstance(p, _Param_Dist) for p in parameters):
            raise ValueError("`parameters` must be a non-empty list of `_Param_Dist` instances!")
        self._parameters = parameters
        self.param_names = [p.name for p in parameters]
        self._nparameters = len(parameters)
        self._nsamples = nsamples
        self._seed = seed
        self.sam_shape = sam_shape
        if nsamples is None:
            self._uniform_samples = None
            self.param_samples = None
        else:
            lhc = scipy.stats.qmc.LatinHypercube(d=self._nparameters, strength=1, seed=seed)
            self._uniform_samples = lhc.random(n=nsamples)                        # (S, D)
            self.param_samples = np.zeros_like(self._uniform_samples)
            for ii, param in enumerate(parameters):
                self.param_samples[:, ii] = param(self._uniform_samples[:, ii])

    @property
    def name(self):
        return self.__class__.__name__

    @property
    def nsamples(self):
        return self._nsamples

    @property
    def nparameters(self):
        return self._nparameters

    @property
    def extrema(self):
        return np.array([p.extrema for p in self._parameters])

    def param_dict(self, samp_num):
        return {name: float(v) for name, v in zip(self.param_names, self.param_samples[samp_num])}

    def normalized_params(self, vals):
        """Uniform [0, 1] values (None: the parameter's default) -> parameter dictionary."""
        vals = np.atleast_1d(vals)
        if len(vals) != self._nparameters:
            raise ValueError(f"{len(vals)=} != {self._nparameters=}")
        return {p.name: (p.default if v is None else float(p(v))) for p, v in zip(self._parameters, vals)}

    def model_for_params(self, params, sam_shape=None, cosmo=None):
        if sam_shape is None:
            sam_shape = self.sam_shape
        settings = dict(self.DEFAULTS)
        for name, value in params.items():
            if name not in settings:
                raise ValueError(f"`{name}` is not a parameter of {self.name}!")
            settings[name] = value
        sam = self._init_sam(sam_shape, settings, cosmo=cosmo)
        hard = self._init_hard(sam, settings)
        return sam, hard

    def model_for_sample_number(self, samp_num, sam_shape=None, cosmo=None):
        return self.model_for_params(self.param_dict(samp_num), sam_shape, cosmo=cosmo)

    @classmethod
    @abc.abstractmethod
    def _init_sam(cls, sam_shape, params, cosmo=None):
        pass

    @classmethod
    @abc.abstractmethod
    def _init_hard(cls, sam, params):
        pass


class _PS_Classic_Phenom(_Param_Space):
    """Classic phenomenological space of the 15-yr analysis (param_spaces_classic.py:9-98)."""

    DEFAULTS = dict(
        hard_time=3.0,          # [Gyr]
        hard_sepa_init=1e4,     # [pc]
        hard_rchar=100.0,       # [pc]
        hard_gamma_inner=-1.0,
        hard_gamma_outer=+2.5,
        gsmf_phi0_log10=-2.77,
        gsmf_phiz=-0.6,
        gsmf_mchar0_log10=11.24,
        gsmf_mcharz=0.11,
        gsmf_alpha0=-1.21,
        gsmf_alphaz=-0.03,
        gpf_frac_norm_allq=0.025,
        gpf_malpha=0.0,
        gpf_qgamma=0.0,
        gpf_zbeta=1.0,
        gpf_max_frac=1.0,
        gmt_norm=0.5,           # [Gyr]
        gmt_malpha=0.0,
        gmt_qgamma=-1.0,
        gmt_zbeta=-0.5,
        mmb_mamp_log10=8.69,
        mmb_plaw=1.10,
        mmb_scatter_dex=0.3,
    )

    @classmethod
    def _init_sam(cls, sam_shape, params, cosmo=None):
        gsmf = sams.GSMF_Schechter(phi0=params['gsmf_phi0_log10'], phiz=params['gsmf_phiz'],
                                   mchar0_log10=params['gsmf_mchar0_log10'], mcharz=params['gsmf_mcharz'],
                                   alpha0=params['gsmf_alpha0'], alphaz=params['gsmf_alphaz'])
        gpf = sams.GPF_Power_Law(frac_norm_allq=params['gpf_frac_norm_allq'], malpha=params['gpf_malpha'],
                                 qgamma=params['gpf_qgamma'], zbeta=params['gpf_zbeta'],
                                 max_frac=params['gpf_max_frac'])
        gmt = sams.GMT_Power_Law(time_norm=params['gmt_norm'] * GYR, malpha=params['gmt_malpha'],
                                 qgamma=params['gmt_qgamma'], zbeta=params['gmt_zbeta'], cosmo=cosmo)
        mmbulge = host_relations.MMBulge_KH2013(mamp_log10=params['mmb_mamp_log10'], mplaw=params['mmb_plaw'],
                                                scatter_dex=params['mmb_scatter_dex'])
        return Semi_Analytic_Model(gsmf=gsmf, gpf=gpf, gmt=gmt, mmbulge=mmbulge, shape=sam_shape, cosmo=cosmo)

    @classmethod
    def _init_hard(cls, sam, params):
        return hardening.Fixed_Time_2PL_SAM(sam, params['hard_time'] * GYR, sepa_init=params['hard_sepa_init'] * PC,
                                            rchar=params['hard_rchar'] * PC,
                                            gamma_inner=params['hard_gamma_inner'],
                                            gamma_outer=params['hard_gamma_outer'])


class PS_Classic_Phenom_Uniform(_PS_Classic_Phenom):
    """Classic uniform phenomenological space (param_spaces_classic.py:100-117)."""

    def __init__(self, nsamples=None, sam_shape=None, seed=None):
        parameters = [
            PD_Uniform("gsmf_phi0_log10", -3.5, -1.5),
            PD_Uniform("gsmf_mchar0_log10", 10.5, 12.5),   # [log10(Msol)]
            PD_Uniform("mmb_mamp_log10", +7.6, +9.0),      # [log10(Msol)]
            PD_Uniform("mmb_scatter_dex", +0.0, +0.9),
            PD_Uniform("hard_time", 0.1, 11.0),            # [Gyr]
            PD_Uniform("hard_gamma_inner", -1.5, +0.0),
        ]
        super().__init__(parameters, nsamples=nsamples, sam_shape=sam_shape, seed=seed)


# ---------------------------------------------------------------------------
# one model (lib_tools.py:714-842)
# ---------------------------------------------------------------------------
def run_model(sam, hard, pta_dur=DEF_PTA_DUR, nfreqs=DEF_NUM_FBINS, nreals=DEF_NUM_REALS, nloudest=DEF_NUM_LOUDEST,
              gwb_flag=True, singles_flag=True, details_flag=False, params_flag=False, *, seed=0):
    """Population and GW signatures of one model; same dictionary keys as the reference.

    ``pta_dur`` is in years.  ``seed`` keys the Philox streams: the loud/background split uses
    ``seed``, the independent Poisson GWB (``gwb_flag``) uses ``seed + 1``.
    """
    from . import engine, _lib
    from . import device as D
    if not any([gwb_flag, details_flag, singles_flag, params_flag]):
        raise RuntimeError(f"No flags set!  {gwb_flag=} {details_flag=} {singles_flag=} {params_flag=}")
    if not isinstance(hard, (hardening.Fixed_Time_2PL_SAM, hardening.Hard_GW)):
        raise RuntimeError(f"`Fixed_Time_2PL_SAM` or `Hard_GW` must be used here!  Not {hard}!")
    data = {}
    fobs_cents, fobs_edges = utils.pta_freqs(dur=pta_dur * YR, num=nfreqs)
    data['fobs_cents'] = fobs_cents
    data['fobs_edges'] = fobs_edges
    number, h2fdf, redz_final = sam._number_h2fdf_device(fobs_edges, hard)
    if details_flag:
        data['static_binary_density'] = sam.static_binary_density
        data['number'] = D.to_host(number).reshape(tuple(n - 1 for n in sam.shape) + (nfreqs,))
        data['redz_final'] = D.to_host(redz_final)
    if singles_flag or params_flag:
        from . import single_sources
        nl = nloudest if singles_flag else 1
        edges = [sam.mtot, sam.mrat, sam.redz, np.asarray(fobs_edges) / 2.0]
        out = single_sources._ss_gws_redz_device(edges, redz_final, number, h2fdf, realize=nreals, loudest=nl,
                                                 params=params_flag, seed=seed, counts=None, loud_mode='sorted',
                                                 cosmo=sam._cosmo)
        sam._check_scatter()
        if params_flag:
            hc_ss, hc_bg, sspar, bgpar = out
            data['sspar'] = sspar
            data['bgpar'] = bgpar
        else:
            hc_ss, hc_bg = out
        if singles_flag:
            data['hc_ss'] = hc_ss
            data['hc_bg'] = hc_bg
    if gwb_flag:
        _, _, hc2, _ = engine.realize(number, h2fdf, int(nreals), 0, _lib.HD_SEL_NONE, seed=seed + 1)
        data['gwb'] = np.sqrt(D.to_host(hc2))
    return data


# ---------------------------------------------------------------------------
# a library (gen_lib.py:50-337 sample loop + combine.py:86-270 file layout)
# ---------------------------------------------------------------------------
def sample_shapes(nsamples, seed, mtot=(41, 91), mrat=(41, 81), redz=(41, 101)):
    """Per-sample grid shapes (edge counts), uniform on the given inclusive ranges from one seed."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(lo, hi + 1, size=nsamples) for lo, hi in (mtot, mrat, redz)], axis=1)


def sim_fname(sims_dir, snum):
    """Per-sample output file of a library run (lib_tools._get_sim_fname: ``sim_{snum:06d}.npz``)."""
    import os
    return os.path.join(sims_dir, f"sim_{int(snum):06d}.npz")


def _sample_done(fname, recreate):
    """True when ``fname`` holds a finished sample to keep (gen_lib.py:283-294): an existing file
    without the 'fail' key, unless ``recreate``."""
    import os
    if recreate or not os.path.exists(fname):
        return False
    try:
        with np.load(fname, allow_pickle=False) as fh:
            return 'fail' not in fh.files
    except (OSError, ValueError):   # an unreadable (e.g. truncated) file is recomputed
        return False


def run_library(pspace, shapes=None, pta_dur=DEF_PTA_DUR, nfreqs=DEF_NUM_FBINS, nreals=DEF_NUM_REALS,
                nloudest=DEF_NUM_LOUDEST, gwb_flag=True, singles_flag=True, params_flag=False, seed=0,
                path=None, samples=None, cosmo=None, log_every=0, sims_dir=None, recreate=False):
    """Run every sample of `pspace` (or the listed sample numbers) and collect the library arrays.

    ``shapes``: (S, 3) grid edge counts per sample, or None for ``pspace.sam_shape``.
    Returns the dictionary of the combined library file and writes it to ``path`` (``.npz``) if
    given.  Sample ``i`` uses the Philox seed ``seed + 2 i``.

    Failures are isolated per sample (gen_lib.py:287-320, combine.py:206-212): an exception in a
    sample's model or run, or a non-finite output, marks that sample failed -- its outputs are NaN,
    its ``sample_params`` row NaN, ``failed[i]`` True and ``fail_messages[i]`` the reason -- and the
    other samples go on.  With ``sims_dir`` every sample is also written to its own file
    (``sim_fname``; a failure as a file with the single key 'fail'), and a rerun resumes: samples
    whose file exists without 'fail' are read back instead of recomputed (all are recomputed with
    ``recreate``).
    """
    import os
    samples = np.arange(pspace.nsamples) if samples is None else np.asarray(samples)
    fobs_cents, fobs_edges = utils.pta_freqs(dur=pta_dur * YR, num=nfreqs)
    S = len(samples)
    lib = {'fobs_cents': fobs_cents, 'fobs_edges': fobs_edges,
           'sample_params': pspace.param_samples[samples].copy()}
    keys = []
    if gwb_flag:
        lib['gwb'] = np.zeros((S, nfreqs, nreals))
        keys.append('gwb')
    if singles_flag:
        lib['hc_ss'] = np.zeros((S, nfreqs, nreals, nloudest))
        lib['hc_bg'] = np.zeros((S, nfreqs, nreals))
        keys += ['hc_ss', 'hc_bg']
    if params_flag:
        lib['sspar'] = np.zeros((S, 4, nfreqs, nreals, nloudest))
        lib['bgpar'] = np.zeros((S, 7, nfreqs, nreals))
        keys += ['sspar', 'bgpar']
    failed = np.zeros(S, dtype=bool)
    messages = [''] * S
    resumed = 0
    if sims_dir is not None:
        os.makedirs(sims_dir, exist_ok=True)
    t0 = time.perf_counter()
    for ii, snum in enumerate(samples):
        fname = None if sims_dir is None else sim_fname(sims_dir, snum)
        if fname is not None and _sample_done(fname, recreate):
            with np.load(fname, allow_pickle=False) as fh:
                for key in keys:
                    lib[key][ii] = fh[key]
            resumed += 1
            continue
        shape = None if shapes is None else tuple(int(v) for v in shapes[ii])
        try:
            sam, hard = pspace.model_for_sample_number(int(snum), sam_shape=shape, cosmo=cosmo)
            data = run_model(sam, hard, pta_dur=pta_dur, nfreqs=nfreqs, nreals=nreals, nloudest=nloudest,
                             gwb_flag=gwb_flag, singles_flag=singles_flag, params_flag=params_flag,
                             seed=seed + 2 * int(snum))
            bad = [key for key in keys if not np.all(np.isfinite(data[key]))]
            if bad:
                raise FloatingPointError(f"non-finite values in {bad}")
        except Exception as err:   # the sample fails, the library goes on
            failed[ii] = True
            messages[ii] = f"{type(err).__name__}: {err}"
            for key in keys:
                lib[key][ii] = np.nan
            if fname is not None:
                np.savez(fname, fail=np.array(messages[ii]))
            continue
        for key in keys:
            lib[key][ii] = data[key]
        if fname is not None:
            out = {key: data[key] for key in keys}
            out['params'] = pspace.param_samples[int(snum)]
            out['param_names'] = np.array(pspace.param_names).astype('S')
            np.savez(fname, **out)
        if log_every and (ii + 1) % log_every == 0:
            dt = time.perf_counter() - t0
            print(f"library: {ii + 1}/{S} samples, {dt:.1f} s ({(ii + 1) / dt:.2f} samples/s)", flush=True)
    lib['sample_params'][failed] = np.nan
    lib['failed'] = failed
    lib['fail_messages'] = np.array(messages)
    lib['resumed'] = np.array(resumed)
    lib['param_names'] = np.array(pspace.param_names).astype('S')
    lib['parameter_space_class_name'] = np.array(pspace.name)
    if path is not None:
        np.savez(path, **lib)
    return lib


def gwb_library(pspace, shapes=None, pta_dur=DEF_PTA_DUR, nfreqs=DEF_NUM_FBINS, nreals=DEF_NUM_REALS, seed=0,
                samples=None, cosmo=None, streams=3, chunk=256, reduce=None, reduce_key='targets', plan_workers=8,
                lookahead=16, log_every=0, path=None, resume=True):
    """The independent Poisson GWB of every sample (``run_model``'s ``gwb_flag`` output alone),
    with samples in flight on several HIP streams (:func:`gwb_for_params`).

    Sample ``i`` uses the Philox seed ``seed + 2 i + 1``, so ``gwb`` equals :func:`run_library`'s
    for the same ``seed``.  ``reduce``: None keeps ``gwb`` (S, F, R); else a function mapping a
    block's (n, F, R) strain to per-sample rows (n, ...), kept under ``reduce_key`` (e.g.
    percentile targets) instead of ``gwb``.  A sample that fails (:func:`gwb_for_params`) has NaN
    outputs and ``sample_params``, ``failed`` True and its reason in ``fail_messages``.

    With ``path`` the library is written there after every block of samples (a checkpoint:
    ``done`` marks the samples finished), and, with ``resume``, an existing file of the same
    samples, seed, frequencies and realizations is continued: only samples not done (never run,
    or failed) are computed.  Returns the library dictionary plus ``samples_per_s`` (this call,
    wall clock, over the samples it computed).
    """
    import os
    samples = np.arange(pspace.nsamples) if samples is None else np.asarray(samples)
    S = samples.size
    fobs_cents, fobs_edges = utils.pta_freqs(dur=pta_dur * YR, num=int(nfreqs))
    params = pspace.param_samples[samples]
    seeds = seed + 2 * samples.astype(np.int64) + 1
    key = 'gwb' if reduce is None else reduce_key
    lib = {'fobs_cents': fobs_cents, 'fobs_edges': fobs_edges, 'sample_params': params.copy(),
           'samples': samples.copy(), 'seed': np.array(seed), 'nreals': np.array(int(nreals)),
           'done': np.zeros(S, dtype=bool), 'failed': np.zeros(S, dtype=bool),
           'fail_messages': np.array([''] * S, dtype=object),
           'param_names': np.array(pspace.param_names).astype('S'),
           'parameter_space_class_name': np.array(pspace.name)}
    if shapes is not None:
        lib['sam_shapes'] = np.asarray(shapes)
    lib[key] = np.full((S, int(nfreqs), int(nreals)), np.nan) if reduce is None else None
    if path is not None and resume and os.path.exists(path):
        with np.load(path, allow_pickle=False) as fh:
            same = (key in fh.files and np.array_equal(fh['samples'], samples) and int(fh['seed']) == int(seed)
                    and int(fh['nreals']) == int(nreals) and np.array_equal(fh['fobs_edges'], fobs_edges))
            if not same:
                raise ValueError(f"{path} holds another library (samples, seed, frequencies or realizations "
                                 "differ); remove it or pass resume=False")
            lib['done'] = fh['done'].copy()
            lib[key] = fh[key].copy()
            lib['fail_messages'] = fh['fail_messages'].astype(object)
    todo = np.flatnonzero(~lib['done'])
    lib['failed'][todo] = False
    lib['fail_messages'][todo] = ''

    def save():
        out = dict(lib)
        out['sample_params'] = np.where(lib['failed'][:, None], np.nan, lib['sample_params'])
        out['fail_messages'] = lib['fail_messages'].astype(str)
        tmp = path + ".tmp.npz"
        np.savez(tmp, **out)
        os.replace(tmp, path)

    def on_block(start, hc, fl, ms):
        rows = todo[start:start + hc.shape[0]]
        vals = hc if reduce is None else np.asarray(reduce(hc))
        if lib[key] is None:
            lib[key] = np.full((S,) + vals.shape[1:], np.nan)
        vals = np.array(vals, dtype=float)
        vals[np.asarray(fl)] = np.nan
        lib[key][rows] = vals
        lib['failed'][rows] = fl
        lib['fail_messages'][rows] = ms
        lib['done'][rows] = ~np.asarray(fl)
        if path is not None:
            save()

    t0 = time.perf_counter()
    info = {'streams': 0}
    if todo.size:
        gwb_for_params(pspace, params[todo], shapes=None if shapes is None else np.asarray(shapes)[todo],
                       fobs_edges=fobs_edges, nreals=nreals, seeds=seeds[todo], cosmo=cosmo, streams=streams,
                       chunk=chunk, plan_workers=plan_workers, lookahead=lookahead, log_every=log_every, info=info,
                       on_block=on_block, collect=False)
    lib['samples_per_s'] = todo.size / (time.perf_counter() - t0) if todo.size else 0.0
    lib['streams_used'] = np.array(info['streams'])
    lib['sample_params'] = np.where(lib['failed'][:, None], np.nan, lib['sample_params'])
    lib['fail_messages'] = lib['fail_messages'].astype(str)
    if path is not None and todo.size == 0:
        save()
    return lib


AUTOTUNE_STREAMS = (2, 3, 4)     # stream counts tried by streams='auto'
AUTOTUNE_PROBE = 16              # rows timed per candidate


def gwb_for_params(pspace, params, shapes=None, fobs_edges=None, nreals=DEF_NUM_REALS, seeds=0, cosmo=None,
                   streams=3, chunk=256, reduce=None, plan_workers=8, lookahead=16, log_every=0, info=None,
                   on_block=None, collect=True):
    """Poisson GWB h_c (n, F, R) for parameter rows ``params`` (n, D) of ``pspace`` (unvaried
    parameters from its ``DEFAULTS``), or ``reduce`` of it per block of samples, concatenated.

    Row ``i`` runs on stream ``i % streams`` (default priority): hardening norms, density and
    scatter, evolution, number, strain and realizations are launched there, and its (F, R) result
    stays in HBM until an asynchronous copy into a pinned host block of ``chunk`` rows, read after
    one synchronisation per block; the host runs at most 2 * ``streams`` rows ahead of the device
    (an event per row), which bounds the pinned staging memory of the uploads.  With per-row
    ``shapes``, the scatter geometry of grids not seen
    before is built by ``plan_workers`` host processes up to ``lookahead`` rows ahead (0: in line,
    when a row needs it).  ``seeds``: one Philox seed per row, or a base seed ``s`` giving row
    ``i`` the seed ``s + i``.

    ``streams='auto'``: the first rows are processed in blocks of AUTOTUNE_PROBE rows, one block
    per candidate in AUTOTUNE_STREAMS, each timed to completion; the fastest count runs the rest
    (rows are never recomputed: a row's result does not depend on its stream).  ``info``, if a
    dict, receives the chosen count, the probe rates, ``failed`` (S,) and ``fail_messages``.

    A row whose model or launch raises, whose scatter leaves bad values, or whose GWB is not finite
    fails alone: its h_c is NaN and ``failed`` marks it; the other rows are unaffected.
    ``on_block(start, hc (n, F, R), failed (n,), messages)`` is called after every block (the
    checkpoint hook of :func:`gwb_library`); with ``collect=False`` nothing else is kept and None
    is returned.
    """
    import torch
    from . import engine, _lib, scatter as _scatter
    from . import device as D
    dev = D.require_device()
    params = np.atleast_2d(np.asarray(params, dtype=float))
    S = params.shape[0]
    if fobs_edges is None:
        fobs_edges = utils.pta_freqs(dur=DEF_PTA_DUR * YR, num=DEF_NUM_FBINS)[1]
    fobs_edges = np.asarray(fobs_edges, dtype=float)
    F, R = fobs_edges.size - 1, int(nreals)
    seeds = np.asarray(seeds, dtype=np.int64)
    if seeds.ndim == 0:
        seeds = int(seeds) + np.arange(S, dtype=np.int64)
    if seeds.shape != (S,):
        raise ValueError(f"{seeds.shape=} must be () or ({S},)")
    if cosmo is None:
        from .cosmology import cosmo as _cos
        cosmo = _cos
    cosmo.device_tables()
    torch.cuda.synchronize()
    if S == 0:
        return np.zeros((0, F, R)) if reduce is None else None

    pool = _scatter.geometry_pool(int(plan_workers)) if (plan_workers and shapes is not None) else None
    pending = {}
    sams_ = {}

    def prefetch(ii):
        # the (host-only) SAM of row ii, and its grid's scatter geometry built by a worker process
        nonlocal pool
        if ii >= S or ii in sams_:
            return
        shape = pspace.sam_shape if shapes is None else tuple(int(v) for v in shapes[ii])
        settings = dict(pspace.DEFAULTS)
        for name, value in zip(pspace.param_names, params[ii]):
            settings[name] = float(value)
        try:
            sam = pspace._init_sam(shape, settings, cosmo=cosmo)
        except Exception as err:      # reported when the row's turn comes
            sams_[ii] = (None, err)
            return
        sams_[ii] = (sam, settings)
        if pool is None or sam._mmbulge._scatter_dex <= 0.0:
            return
        key = _scatter.plan_key(sam.mtot, sam.mrat)
        if key in pending or _scatter.has_plan(sam.mtot, sam.mrat):
            return
        try:
            pending[key] = pool.submit(_scatter.build_geometry, sam.mtot, sam.mrat, 4, 1)
        except Exception:           # a broken pool: geometry is then built in line when needed
            pool = None

    auto = isinstance(streams, str) and streams == 'auto'
    cands = list(AUTOTUNE_STREAMS) if auto and S >= 4 * AUTOTUNE_PROBE * len(AUTOTUNE_STREAMS) else []
    nstream = max(cands) if cands else (3 if auto else max(1, int(streams)))
    streams_ = [torch.cuda.Stream(device=dev) for _ in range(max([nstream] + cands))]
    block = torch.empty((min(max(chunk, AUTOTUNE_PROBE), S), F, R), dtype=torch.float64, pin_memory=True)
    out = (np.zeros((S, F, R)) if reduce is None else []) if collect else None
    inflight = collections.deque()      # one event per launched row: the host stays <= depth rows ahead
    probe_rates = {}
    failed = np.zeros(S, dtype=bool)
    messages = [''] * S
    checks = [None] * block.shape[0]     # per row of a block: the scatter's status tensor
    t0 = time.perf_counter()
    try:
        for ii in range(min(lookahead, S)):
            prefetch(ii)
        start = 0
        while start < S:
            if cands:                           # autotuning block: the next candidate stream count
                nstream = cands[len(probe_rates)]
                n = AUTOTUNE_PROBE
                tb = time.perf_counter()
            else:
                n = min(chunk, S - start)
            depth = 2 * nstream
            for j in range(n):
                ii = start + j
                prefetch(ii + lookahead)
                sam, settings = sams_.pop(ii)
                if sam is None:               # the model itself could not be built
                    failed[ii] = True
                    messages[ii] = f"{type(settings).__name__}: {settings}"
                    block[j].fill_(float("nan"))
                    checks[j] = None
                    continue
                fut = pending.pop(_scatter.plan_key(sam.mtot, sam.mrat), None)
                if fut is not None:
                    try:
                        geom = fut.result()
                    except Exception:       # a broken worker pool (e.g. no importable __main__): in line
                        geom, pool = None, None
                    _scatter.put_plan(_scatter.ScatterPlan(sam.mtot, sam.mrat, geom=geom), sam.mtot, sam.mrat)
                try:
                    with torch.cuda.stream(streams_[ii % nstream]):
                        hard = pspace._init_hard(sam, settings)
                        number, h2fdf, _ = sam._number_h2fdf_device(fobs_edges, hard)
                        _, _, hc2, _ = engine.realize(number, h2fdf, R, 0, _lib.HD_SEL_NONE, seed=int(seeds[ii]))
                        block[j].copy_(hc2, non_blocking=True)
                        ev = torch.cuda.Event()
                        ev.record()
                    checks[j] = sam._scatter_iters
                except Exception as err:   # this row fails, the others go on (gen_lib.py:307-316)
                    failed[ii] = True
                    messages[ii] = f"{type(err).__name__}: {err}"
                    block[j].fill_(float("nan"))
                    checks[j] = None
                    continue
                inflight.append(ev)
                if len(inflight) > depth:
                    inflight.popleft().synchronize()
            torch.cuda.synchronize()
            inflight.clear()
            if cands:
                probe_rates[nstream] = n / (time.perf_counter() - tb)
                if len(probe_rates) == len(cands):
                    nstream = max(probe_rates, key=probe_rates.get)
                    cands = []
            hc = np.sqrt(block[:n].numpy())
            for j in range(n):   # the scatter's nearest-node fill (sam.py:1377-1383) and finite outputs
                ii = start + j
                if failed[ii]:
                    continue
                it = checks[j]
                bad_fill = 0 if it is None else int(it[-1].item())
                if bad_fill or not np.all(np.isfinite(hc[j])):
                    failed[ii] = True
                    messages[ii] = ("ValueError: M-Mbulge scatter left bad values" if bad_fill
                                    else "FloatingPointError: non-finite GWB")
                    hc[j] = np.nan
            if on_block is not None:
                on_block(start, hc, failed[start:start + n], messages[start:start + n])
            if not collect:
                pass
            elif reduce is None:
                out[start:start + n] = hc
            else:
                out.append(np.asarray(reduce(hc)))
            start += n
            if log_every and (start // log_every) != ((start - n) // log_every):
                dt = time.perf_counter() - t0
                print(f"gwb_for_params: {start}/{S} samples, {dt:.1f} s ({start / dt:.2f} samples/s)", flush=True)
    finally:
        for fut in pending.values():
            fut.cancel()
    if isinstance(info, dict):
        info.update(streams=nstream, probe_rates=probe_rates, failed=failed, fail_messages=messages)
    if not collect:
        return None
    return out if reduce is None else np.concatenate(out, axis=0)


def eccen_library(pspace, eccen_inits, samples=None, shape=(61, 51, 61), pta_dur=DEF_PTA_DUR, nfreqs=DEF_NUM_FBINS,
                  nharms=100, sepa_init=1.0 * PC, nsteps=123, cosmo=None, log_every=0):
    """Eccentric-SAM spectra of parameter samples x initial eccentricities (as the reference's
    eccentric library script, scripts/gen_lib_sams_eccen.py:744-790, with the deterministic
    harmonic sum of cyutils.pyx:361-597 in place of its Poisson variant).

    Per sample: the static binary density (with M-Mbulge scatter) on the device once; per initial
    eccentricity: the GW-driven e(a) from ``sepa_init`` over ``nsteps`` (sams/sam.py:1235-1288)
    and the (F, nharms) harmonic sum at the frequencies i/T.  Returns a dictionary with ``hc``
    (S, E0, F) = sqrt of the harmonic-summed strain squared, ``sample_params`` (S, D),
    ``eccen_init`` (E0,), ``fobs`` (F,) and ``samples_per_s``.
    """
    import torch
    from . import engine, gravwaves
    from .sams.sam import evolve_eccen_uniform_single
    samples = np.arange(pspace.nsamples) if samples is None else np.asarray(samples)
    e0s = np.asarray(eccen_inits, dtype=float)
    fobs = utils.pta_freqs(dur=pta_dur * YR, num=int(nfreqs))[0]
    if cosmo is None:
        from .cosmology import cosmo as _cos
        cosmo = _cos
    hc = np.zeros((samples.size, e0s.size, fobs.size))
    t0 = time.perf_counter()
    for ii, snum in enumerate(samples):
        sam, _ = pspace.model_for_sample_number(int(snum), sam_shape=tuple(shape), cosmo=cosmo)
        outs = []
        for e0 in e0s:
            sepa, ecc = evolve_eccen_uniform_single(sam, float(e0), sepa_init, int(nsteps))
            g = engine.gwb_eccen(*gravwaves._eccen_inputs(fobs, sam, sepa, ecc), int(nharms))
            outs.append(torch.sqrt(g.sum(dim=1)))
        hc[ii] = torch.stack(outs).cpu().numpy()
        if log_every and (ii + 1) % log_every == 0:
            dt = time.perf_counter() - t0
            print(f"eccen_library: {ii + 1}/{samples.size} samples, {dt:.1f} s", flush=True)
    return {'hc': hc, 'sample_params': pspace.param_samples[samples].copy(), 'eccen_init': e0s, 'fobs': fobs,
            'param_names': np.array(pspace.param_names).astype('S'),
            'samples_per_s': samples.size / (time.perf_counter() - t0)}


def load_library(path):
    """Read a library written by :func:`run_library` (no pickled objects)."""
    with np.load(path, allow_pickle=False) as fh:
        return {k: fh[k] for k in fh.files}
