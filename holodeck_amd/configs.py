"""Model configurations of the benchmark workloads, as plain specification dicts,
and builders of the corresponding SAM and hardening objects."""

from holodeck_amd.constants import MSOL, PC, GYR, YR
from holodeck_amd import utils


def spec_config1(shape=(61, 51, 41), nfreq=20):
    """Plain SAM: flat LCDM (H0=67.7, Om=0.31); GSMF Schechter; GPF/GMT power laws; KH2013 M-Mbulge; GW-only."""
    return dict(
        shape=shape, mtot=(10 ** 4.5 * MSOL, 1e11 * MSOL), mrat=(1e-3, 1.0), redz=(1e-3, 10.0),
        cosmo=dict(h=0.677, Om0=0.31),
        gsmf=dict(phi0=-2.77, phiz=-0.27, mchar0_log10=10.84, mcharz=0.11, alpha0=-1.24, alphaz=-0.03),
        gpf=dict(frac_norm_allq=0.025, zbeta=0.8), gmt=dict(time_norm=0.55 * GYR, zbeta=-0.5), gmr=None,
        mmbulge=dict(mamp_log10=8.69, mplaw=1.17, scatter_dex=0.28),
        hard=dict(kind='gw'), dur=16.0 * YR, nfreq=nfreq,
    )


def spec_config2(shape=(91, 81, 101), nfreq=40):
    """Full-size SAM: 91x81x101, 40 frequencies over 16.03 yr, Fixed_Time_2PL_SAM(1 Gyr, -1, +2.5, 100 pc, 1e4 pc)."""
    s = spec_config1(shape=shape, nfreq=nfreq)
    s.update(mtot=(1e4 * MSOL, 1e12 * MSOL), dur=16.03 * YR,
             hard=dict(kind='2pl', time=1.0 * GYR, sepa_init=1e4 * PC, rchar=100.0 * PC, gamma_inner=-1.0,
                       gamma_outer=2.5, num_steps=300))
    return s


def fobs_gw_edges(spec):
    return utils.pta_freqs(spec['dur'], spec['nfreq'])[1]


def make_sam(spec, scatter=None, cosmo=None):
    import holodeck_amd as holo
    cos = cosmo if cosmo is not None else holo.Cosmology(h=spec["cosmo"]["h"], Om0=spec["cosmo"]["Om0"])
    gsmf = holo.sams.GSMF_Schechter(**spec['gsmf'])
    kw = {}
    if spec.get('gmr') is not None:
        kw['gmr'] = holo.sams.GMR_Illustris(**spec['gmr'])
        if spec.get('gmt') is not None:
            kw['gmt'] = holo.sams.GMT_Power_Law(cosmo=cos, **spec['gmt'])
    else:
        kw['gpf'] = holo.sams.GPF_Power_Law(**spec['gpf'])
        kw['gmt'] = holo.sams.GMT_Power_Law(cosmo=cos, **spec['gmt'])
    mm = dict(spec['mmbulge'])
    if scatter is not None:
        mm['scatter_dex'] = scatter
    mmb = holo.host_relations.MMBulge_KH2013(**mm)
    M, Q, Z = spec['shape']
    sam = holo.sams.Semi_Analytic_Model(mtot=(*spec['mtot'], M), mrat=(*spec['mrat'], Q), redz=(*spec['redz'], Z),
                                        gsmf=gsmf, mmbulge=mmb, cosmo=cos, **kw)
    return sam


def make_hard(spec, sam):
    import holodeck_amd as holo
    h = spec['hard']
    if h['kind'] == 'gw':
        return holo.hardening.Hard_GW()
    return holo.hardening.Fixed_Time_2PL_SAM(sam, h['time'], sepa_init=h['sepa_init'], rchar=h['rchar'],
                                             gamma_inner=h['gamma_inner'], gamma_outer=h['gamma_outer'],
                                             num_steps=h['num_steps'])
