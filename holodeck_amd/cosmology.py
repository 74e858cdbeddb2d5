"""Flat LambdaCDM cosmology with FP64 device tables.

Replaces the `holo.cosmo` instance the reference builds from cosmopy
(holodeck/__init__.py:81-86; WMAP9 parameters at holodeck/__init__.py:48-53).
On the hot path the reference uses:

* the interpolation grids ``_grid_z`` (decreasing, ending at z=0),
  ``_grid_age`` [s] and ``_grid_dcom`` [cm] (sam_cyutils.pyx:476, 494);
* ``age``, ``tage_to_z`` (via utils.redz_after, utils.py:1772-1808), ``dtdz``
  (sam.py:347) and ``comoving_distance`` (gravwaves.py:713-714).

Here the grids have ``size`` nodes uniform in log10(1+z) from ``zmax`` to 0
(the reference's node placement lives in the absent cosmopy package), and all
quantities are evaluated on the device in FP64: age, dt/dz and t->z in closed
form, the comoving distance from a 4096-node cubic Hermite table of the exact
integral.  Radiation is neglected, as in astropy's FlatLambdaCDM with the
default Tcmb0=0.
"""
import numpy as np

from . import _lib
from .device import torch
from . import device as D
from .constants import MPC, SPLC, GYR

DCOM_TABLE_SIZE = 4096


class Cosmology:
    """Flat LambdaCDM: h = H0/(100 km/s/Mpc), Om0 matter density, Ob0 (unused by the path)."""

    def __init__(self, h=0.6933, Om0=0.2880, Ob0=0.0472, size=200, zmax=20.0):
        if not (0.0 < Om0 < 1.0) or not (h > 0.0):
            raise ValueError(f"invalid cosmology h={h} Om0={Om0}")
        if size < 4:
            raise ValueError("size must be >= 4")
        self.h = float(h)
        self.Om0 = float(Om0)
        self.Ob0 = float(Ob0)
        self.Ode0 = 1.0 - self.Om0
        self.H0 = self.h * 100.0 * 1.0e5 / MPC          #: [1/s]
        self.hubble_time = 1.0 / self.H0                  #: [s]
        self.hubble_distance = SPLC / self.H0             #: [cm]
        self.size = int(size)
        self.zmax = float(zmax)
        xx = np.linspace(np.log10(1.0 + self.zmax), 0.0, self.size)
        zz = 10.0 ** xx - 1.0
        zz[-1] = 0.0
        self._grid_z = zz                                  #: decreasing redshift nodes
        self._dev = {}
        self._grid_age_h = None
        self._grid_dcom_h = None

    # ---- device tables ----
    def device_tables(self):
        """dict of device tensors: dc_tab, dc_der (Hermite d_c table), grid_z, grid_age, grid_dcom."""
        dev = D.require_device()
        key = str(dev)
        tabs = self._dev.get(key)
        if tabs is not None:
            return tabs
        n = DCOM_TABLE_SIZE
        dc_tab = D.empty(n + 1)
        dc_der = D.empty(n + 1)
        st = D.stream_ptr()
        _lib.call("hd_cosmo_build_dcom_table", self.Om0, n, D.ptr(dc_tab), D.ptr(dc_der), st)
        gz = D.to_dev(self._grid_z)
        gage = D.empty(self.size)
        gdc = D.empty(self.size)
        _lib.call("hd_cosmo_eval", self.Om0, self.h, D.ptr(dc_tab), D.ptr(dc_der), n, D.ptr(gz), self.size,
                  D.ptr(gage), D.ptr(gdc), None, st)
        tabs = dict(dc_tab=dc_tab, dc_der=dc_der, n_tab=n, grid_z=gz, grid_age=gage, grid_dcom=gdc)
        # built once per device and then read from any stream: finish the build here
        torch.cuda.current_stream().synchronize()
        self._dev[key] = tabs
        return tabs

    def _eval(self, z, which):
        t = self.device_tables()
        zz = np.atleast_1d(np.asarray(z, dtype=float))
        zd = D.to_dev(zz.ravel())
        out = D.empty(zz.size)
        args = [None, None, None]
        args[which] = D.ptr(out)
        _lib.call("hd_cosmo_eval", self.Om0, self.h, D.ptr(t["dc_tab"]), D.ptr(t["dc_der"]), t["n_tab"],
                  D.ptr(zd), zz.size, *args, D.stream_ptr())
        res = D.to_host(out).reshape(zz.shape)
        return res if np.ndim(z) else float(res[0])

    # ---- reference-named quantities (device-evaluated) ----
    def age(self, z):
        """Age of the universe at z [s]."""
        return self._eval(z, 0)

    def comoving_distance(self, z):
        """Line-of-sight comoving distance to z [cm]."""
        return self._eval(z, 1)

    def dtdz(self, z):
        """|dt/dz| = 1/((1+z) H(z)) [s]."""
        return self._eval(z, 2)

    @property
    def _grid_age(self):
        if self._grid_age_h is None:
            self._grid_age_h = D.to_host(self.device_tables()["grid_age"])
        return self._grid_age_h

    @property
    def _grid_dcom(self):
        if self._grid_dcom_h is None:
            self._grid_dcom_h = D.to_host(self.device_tables()["grid_dcom"])
        return self._grid_dcom_h

    def age_universe_gyr(self):
        return float(self.hubble_time * (2.0 / 3.0) / np.sqrt(self.Ode0)
                     * np.arcsinh(np.sqrt(self.Ode0 / self.Om0)) / GYR)

    def __repr__(self):
        return f"Cosmology(h={self.h}, Om0={self.Om0}, Ob0={self.Ob0}, size={self.size})"


class Parameters:
    """WMAP9 parameters of the reference (holodeck/__init__.py:48-53)."""
    Omega0 = 0.2880
    OmegaBaryon = 0.0472
    HubbleParam = 0.6933


#: default cosmology instance, as the reference's ``holo.cosmo``
cosmo = Cosmology(h=Parameters.HubbleParam, Om0=Parameters.Omega0, Ob0=Parameters.OmegaBaryon, size=200)
