"""CPU-only checks of the C-ABI library and the host logic (no GPU compute)."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "holohip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from holodeck_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if " T " in line)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None
    # every declared entry point has a ctypes signature
    assert not [n for n in names if n not in _lib.SIGNATURES]


def test_philox_known_answers():
    """Random123 philox4x32_10 known-answer vectors (kat_vectors)."""
    from holodeck_amd import _lib
    assert _lib.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert _lib.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                       0x6d5451fd]
    assert _lib.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_errors_are_reported_without_gpu():
    """Argument validation happens before any device work and reports through hd_last_error."""
    from holodeck_amd import _lib
    with pytest.raises(_lib.HoloHipError, match="bad arguments"):
        _lib.call("hd_find_2pwl_norm", 1.0, None, None, 5, 1.0, 1.0, -1.0, 2.5, 300, None, None, None)
    with pytest.raises(_lib.HoloHipError, match="L=99"):
        _lib.call("hd_realize_counts", 1, 1, None, 10, 2, 3, 99, 0, 0, None, 0, 1, 1, 1, None, None, 0, None)


def test_launch_timing_switch_without_gpu():
    """Timing records nothing until a timed launch runs; enabling clears, reading an unknown name gives zero."""
    from holodeck_amd import _lib
    _lib.timing_enable(True)
    assert _lib.timing_read("hd_realize") == (0.0, 0)
    assert _lib.timing_read("no_such_kernel") == (0.0, 0)
    _lib.timing_enable(False)
    with pytest.raises(_lib.HoloHipError, match="null argument"):
        _lib.call("hd_timing_read", b"hd_realize", None, None)


def test_pta_freqs_and_midpoints():
    from holodeck_amd import utils
    from holodeck_amd.constants import YR
    cents, edges = utils.pta_freqs(16.03 * YR, 40)
    np.testing.assert_allclose(cents, np.arange(1, 41) / (16.03 * YR))
    np.testing.assert_allclose(utils.midpoints(edges), cents, rtol=1e-14)
    assert utils.isinteger(5) and not utils.isinteger(False.__class__) and not utils.isinteger(5.0)


def test_component_conversions_match_reference_definitions():
    from holodeck_amd import sams, host_relations
    from holodeck_amd.constants import MSOL, GYR
    gpf = sams.GPF_Power_Law(frac_norm_allq=0.025)
    assert np.isclose(gpf._frac_norm, 0.025 / 0.75)            # components.py:531-539
    gmt = sams.GMT_Power_Law(cosmo=type("C", (), {"h": 0.5})())
    assert np.isclose(gmt._mref, 1e11 * MSOL * 0.8)             # components.py:646
    gmr = sams.GMR_Illustris()
    assert np.isclose(gmr._norm0, 10 ** -2.2287 / GYR)
    mm = host_relations.MMBulge_KH2013()
    assert np.isclose(mm._mamp, MSOL * 10 ** 8.69) and mm._mplaw == 1.17 and mm._scatter_dex == 0.28
    assert mm._bulge_frac.bulge_frac() == 0.615
    g = sams.GSMF_Schechter(mchar0_log10=10.84, mcharz=0.11)
    assert np.isclose(g._mchar0, MSOL * 10 ** 10.84) and g._mcharz == 0.11


def test_density_param_layout_matches_kernel_enum():
    """The packed parameter vector follows the index enum of density_par.h."""
    from holodeck_amd import sams, host_relations, cosmology
    from holodeck_amd.sams.components import density_params
    hdr = open(os.path.join(ROOT, "holodeck_amd", "csrc", "density_par.h")).read()
    names = re.findall(r"^\s+(DP_[A-Z0-9_]+)", hdr, flags=re.M)
    pos = {n: i for i, n in enumerate(names)}
    c = cosmology.Cosmology(h=0.7, Om0=0.3)
    p = density_params(sams.GSMF_Schechter(), sams.GPF_Power_Law(), sams.GMT_Power_Law(cosmo=c), None,
                       host_relations.MMBulge_KH2013(), c)
    assert p[pos['DP_GSMF_PHI0']] == -2.77 and p[pos['DP_USE_GMR']] == 0 and p[pos['DP_HAS_GMT']] == 1
    assert p[pos['DP_MMB_PLAW']] == 1.17 and p[pos['DP_COSMO_OM0']] == 0.3 and p[pos['DP_COSMO_H']] == 0.7
    assert p[pos['DP_GPF_MAXFRAC']] == 1.0 and p[pos['DP_MMB_BFRAC']] == 0.615
    p2 = density_params(sams.GSMF_Schechter(), None, None, sams.GMR_Illustris(), host_relations.MMBulge_KH2013(), c)
    assert p2[pos['DP_USE_GMR']] == 1 and p2[pos['DP_HAS_GMT']] == 0
    assert p2[pos['DP_GMR_MREF']] == sams.GMR_Illustris()._mref


@pytest.mark.parametrize("shape", [(17, 15), (41, 81), (91, 81)])
def test_level_schedule_and_transforms_on_the_host(shape):
    """The scatter geometry's host part: hd_ct_levels gives every node one more than the highest
    level of its lower-indexed neighbours (a direct Python loop over scipy's adjacency), so each
    level only reads finished lower neighbours; the closed-form simplex transforms agree with
    scipy.spatial.Delaunay.transform to 1e-12 relative."""
    from scipy.spatial import Delaunay
    from holodeck_amd import scatter
    from holodeck_amd.constants import MSOL
    M, Q = shape
    mtot = np.logspace(np.log10(1e4 * MSOL), np.log10(1e12 * MSOL), M)
    mrat = np.logspace(-3, 0, Q)
    g = scatter.build_geometry(mtot, mrat)
    h = g['host']
    ip, ix, N = h['nbr_ptr'], h['nbr_idx'], g['N']
    want = np.zeros(N, dtype=np.int64)
    for p in range(N):
        nb = ix[ip[p]:ip[p + 1]]
        lo = nb[nb < p]
        want[p] = want[lo].max() + 1 if lo.size else 0
    got = np.zeros(N, dtype=np.int64)
    lp, li = h['levels_ptr'], h['levels_idx']
    for L in range(lp.size - 1):
        got[li[lp[L]:lp[L + 1]]] = L
    assert np.array_equal(got, want)
    for L in range(lp.size - 1):   # a level's nodes have all lower neighbours in earlier levels
        for p in li[lp[L]:lp[L + 1]]:
            nb = ix[ip[p]:ip[p + 1]]
            assert np.all(got[nb[nb < p]] < L)
    T = Delaunay(h['pts']).transform
    assert np.allclose(h['transform'], T, rtol=1e-12, atol=1e-12 * np.abs(T).max())


def test_poisson_window_sizing_on_the_host():
    """The CDF tables of the realization kernels, sized on the host by the same function: the
    support drops less than 2^-56 below its first count kb (scipy's Poisson CDF), the table covers
    lam -/+ 4 sqrt(lam) (or reaches the support's ends), at most `cap` entries, and no table (n = 0)
    when it is too wide or the support's end would overflow an int; the hook rejects such a table
    before launching anything."""
    from scipy.stats import poisson
    from holodeck_amd import _lib
    cap = 5120
    for lam in (1e-3, 0.7, 1.0, 6.3, 143.9, 2.5e4, 1.2e5, 2.0e5, 4.0e5):
        kb, kt, n = _lib.poisson_window(lam, cap)
        if n == 0:
            assert lam > 3.9e5     # tables grow like 8 sqrt(lam): wider than 5120 entries above ~4e5
            continue
        sd = np.sqrt(lam)
        assert 0 < n <= cap and 0 <= kb <= kt
        assert poisson.cdf(kb - 1, lam) < 2.0 ** -56
        assert kt == kb or kt + 1 <= lam - 4.0 * sd
        assert kt + n - 1 >= lam + 4.0 * sd + 1.0 or poisson.sf(kt + n - 1, lam) < 2.0 ** -56
        assert n <= 8.0 * sd + 5.0 or kt == kb
    for lam in (2.0 ** 31, 1e12, 1e30, 1e35):   # rounding would collapse these supports to one node
        assert _lib.poisson_window(lam, cap) == (0, 0, 0)
    assert _lib.poisson_window(1.0, 8000) == (0, 0, 7)
    with pytest.raises(_lib.HoloHipError, match="exceeds cap"):
        _lib.call("hd_poisson_cdf_table", 1e12, 8000, 1, 1, None)
    with pytest.raises(_lib.HoloHipError, match="bad arguments"):
        _lib.call("hd_poisson_cdf_table", 1.0, 8001, 1, 1, None)
