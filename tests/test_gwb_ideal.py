"""Idealised GWB (gravwaves.py:619-663, sam.py:837-870): the oracle pinned by closed forms (CPU)
and the HIP kernels against it (GPU)."""
import numpy as np
import pytest

import oracle.ideal as OI


def _grid(seed=0):
    rng = np.random.default_rng(seed)
    mtot = np.logspace(40, 45, 9)[:, None, None]
    mrat = np.logspace(-2, 0, 7)[None, :, None]
    redz = np.logspace(-2, 0.7, 6)[None, None, :]
    ndens = 10 ** rng.uniform(-80, -75, (9, 7, 6))
    return ndens, mtot, mrat, redz


def test_constant_integrand_gives_the_closed_form():
    """A density that makes the integrand constant integrates exactly: h_c^2 = 4 pi/(3 c^2)
    (pi f)^(-4/3) c0 * range(log10 M) * range(q) * range(z)."""
    _, mtot, mrat, redz = _grid()
    c0 = 3.7e-20
    mc = np.power(OI.NWTG * OI.chirp_mass_mtmr(mtot, mrat), 5 / 3)
    ndens = c0 / (mc * np.power(1 + redz, -1 / 3)) * np.ones((9, 7, 6))
    f = np.array([1e-9, 3e-9, 1e-8])
    vol = np.ptp(np.log10(mtot)) * np.ptp(mrat) * np.ptp(redz)
    want = np.sqrt(4 * np.pi / (3 * OI.SPLC ** 2) * (np.pi * f) ** (-4 / 3) * c0 * vol)
    assert np.allclose(OI.gwb_ideal(f, ndens, mtot, mrat, redz), want, rtol=1e-12)


def test_spectral_index_and_per_node_redshift_path():
    """h_c ~ f^(-2/3) exactly; a per-node redshift array equal at every (M, q) integrates as the
    1-D edges do; nodes at z <= 0 (the -1 sentinel) add nothing and leave the result finite."""
    ndens, mtot, mrat, redz = _grid()
    f = np.array([1e-9, 2e-9, 8e-9])
    a = OI.gwb_ideal(f, ndens, mtot, mrat, redz)
    assert np.allclose(a * f ** (2 / 3), a[0] * f[0] ** (2 / 3), rtol=1e-13)
    b = OI.gwb_ideal(f, ndens, mtot, mrat, redz * np.ones_like(ndens))
    assert np.allclose(a, b, rtol=1e-13)
    z3 = redz * np.ones_like(ndens)
    z3[:, :, 0] = -1.0
    with np.errstate(divide='ignore'):
        c = OI.gwb_ideal(f, ndens, mtot, mrat, z3)
    assert np.all(np.isfinite(c)) and np.all(c > 0)
    cells = OI.gwb_ideal(f, ndens, mtot, mrat, redz, sum=False)
    assert cells.shape == (8, 6, 5, 3) and np.allclose(np.sqrt((cells ** 2).sum(axis=(0, 1, 2))), a, rtol=1e-12)


@pytest.mark.gpu
def test_gwb_ideal_matches_oracle(gpu):
    from holodeck_amd import gravwaves
    print("[gwb ideal]", flush=True)
    ndens, mtot, mrat, redz = _grid(3)
    f = np.array([1e-9, 2e-9, 8e-9, 3e-8])
    z3 = redz * (1 + 0.1 * np.random.default_rng(1).random(ndens.shape))
    z3[0, 0, 0] = -1.0
    for rz in (redz, z3):
        for s in (True, False):
            got = gravwaves.gwb_ideal(f, ndens, mtot, mrat, rz, True, sum=s)
            ref = OI.gwb_ideal(f, ndens, mtot, mrat, rz, sum=s)
            assert got.shape == ref.shape and np.allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_sam_gwb_ideal_uses_the_post_merger_redshifts(gpu):
    """Semi_Analytic_Model.gwb_ideal on the device density: with a GMT the redshifts are z', as
    the reference's default, and with redz_prime=False the grid's redshift edges."""
    from holodeck_amd import configs
    from holodeck_amd.constants import MPC
    print("[sam gwb ideal]", flush=True)
    spec = configs.spec_config1(shape=(21, 17, 15), nfreq=6)
    sam = configs.make_sam(spec)
    f = np.arange(1, 7) / (16.0 * 3.15576e7)
    got = sam.gwb_ideal(f)
    ref = OI.gwb_ideal(f, sam.static_binary_density / MPC ** 3, sam.mtot[:, None, None], sam.mrat[None, :, None],
                       sam._redz_prime)
    assert np.allclose(got, ref, rtol=1e-10, atol=0)
    plain = sam.gwb_ideal(f, redz_prime=False)
    ref2 = OI.gwb_ideal(f, sam.static_binary_density / MPC ** 3, sam.mtot[:, None, None], sam.mrat[None, :, None],
                        sam.redz[None, None, :])
    assert np.allclose(plain, ref2, rtol=1e-10, atol=0)
