"""Eruption source term of the standard single-column sources: the host
side of the reference's Source module (src/Source.f90) over the library's
C ABI.  `Source` holds the eruptive pulses, builds the normalized source
column and the mass flux rates once (Calc_Normalized_SourceCol,
EruptivePulse_MassFluxRate), and per time step finds the active pulses
(CheckEruptivePulses), forms SourceNodeFlux (TephraSourceNodes) and inserts
it with the device column add (src/Ash3d.F90:414-431).  The umbrella
helpers at the end are the host side of src/Source_Umbrella.f90 (stencil,
levels, radial winds, 3x3 patch flux, inserted volume); the device side is
Transport.umbrella_source_add / umbrella_winds_add.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import transport as _tr

SOURCE_TYPES = {"suzuki": 0, "umbrella": 1, "umbrella_air": 2, "line": 3, "point": 4, "profile": 5, "other": 6}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


SourceDesc = _tr.SourceDesc


def _arr(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Source:
    """Eruptive pulses of one vent.

    start_time, duration (h), volume (km3 DRE), plume_height (km) per pulse;
    s_lb (nz+1), s_cc, ds (nz): the vertical s grid (km); bin_mass: the
    fraction of the mass in each grain bin (n_gs_max values).  Profile
    sources take prof_dz (per pulse, km) and prof_volume (pulses x points,
    km3 per point)."""

    def __init__(self, source_type, start_time, duration, volume, plume_height, s_lb, s_cc, ds, bin_mass,
                 suzuki_a=12.0, magma_density=2500.0, z_volcano=0.0, zscaling_id=0, zsurf_vent=0.0, ztop=0.0,
                 prof_dz=None, prof_volume=None):
        self.lib = _tr.load()
        self.type = SOURCE_TYPES[source_type] if isinstance(source_type, str) else int(source_type)
        self._keep = {}
        k = self._keep
        k["start"], k["dur"], k["vol"], k["ph"] = (_arr(v) for v in (start_time, duration, volume, plume_height))
        k["s_lb"], k["s_cc"], k["ds"], k["bm"] = (_arr(v) for v in (s_lb, s_cc, ds, bin_mass))
        ne = len(k["start"])
        self.nz = len(k["s_cc"])
        assert len(k["s_lb"]) == self.nz + 1 and len(k["ds"]) == self.nz
        self.n_gs_max = len(k["bm"])
        d = SourceDesc()
        d.source_type, d.neruptions = self.type, ne
        d.start_time, d.duration = k["start"].ctypes.data_as(_dp), k["dur"].ctypes.data_as(_dp)
        d.volume, d.plume_height = k["vol"].ctypes.data_as(_dp), k["ph"].ctypes.data_as(_dp)
        d.suzuki_a, d.magma_density, d.z_volcano = float(suzuki_a), float(magma_density), float(z_volcano)
        d.zscaling_id, d.zsurf_vent, d.ztop = int(zscaling_id), float(zsurf_vent), float(ztop)
        d.s_lb, d.s_cc, d.ds = (k[n].ctypes.data_as(_dp) for n in ("s_lb", "s_cc", "ds"))
        d.bin_mass = k["bm"].ctypes.data_as(_dp)
        if prof_volume is not None:
            pv = np.asfortranarray(prof_volume, dtype=np.float64)
            k["pv"] = pv
            k["pdz"] = _arr(prof_dz)
            k["pnz"] = np.full(ne, pv.shape[1], dtype=np.int32)
            d.prof_dz, d.prof_nz = k["pdz"].ctypes.data_as(_dp), k["pnz"].ctypes.data_as(_ip)
            d.prof_volume, d.prof_ld = pv.ctypes.data_as(_dp), ne
        self.desc = d
        self.norm = np.zeros((ne, self.nz), order="F")
        self.mass_flux_rate = np.zeros(ne)
        self.end_time = np.zeros(ne)
        self._check(self.lib.ash3d_tr_source_column(ctypes.byref(d), self.nz, self.norm.ctypes.data_as(_dp),
                                                    self.mass_flux_rate.ctypes.data_as(_dp),
                                                    self.end_time.ctypes.data_as(_dp)))
        self.ieruption = ctypes.c_int32(1)
        self.jeruption = ctypes.c_int32(1)
        self.frac = np.zeros(ne)

    def _check(self, rc):
        if rc:
            raise _tr.Ash3dError(rc, self.lib.ash3d_tr_last_error().decode())

    def pulses(self, time: float, dt: float) -> bool:
        """CheckEruptivePulses for [time, time+dt]; fractions in self.frac."""
        in_dt = ctypes.c_int32(0)
        self._check(self.lib.ash3d_tr_eruptive_pulses(ctypes.byref(self.desc), self.end_time.ctypes.data_as(_dp),
                                                      float(time), float(dt), ctypes.byref(self.ieruption),
                                                      ctypes.byref(self.jeruption), self.frac.ctypes.data_as(_dp),
                                                      ctypes.byref(in_dt)))
        return bool(in_dt.value)

    def nodes(self, kappa_col, ns: int) -> np.ndarray:
        """TephraSourceNodes: SourceNodeFlux(1:nz+1, 1:ns) (kg/km3/hr)."""
        kap = _arr(kappa_col)
        assert kap.shape == (self.nz,)
        flux = np.zeros((self.nz + 1, ns), order="F")
        self._check(self.lib.ash3d_tr_source_nodes(ctypes.byref(self.desc), self.nz, ns, self.n_gs_max,
                                                   self.norm.ctypes.data_as(_dp),
                                                   self.mass_flux_rate.ctypes.data_as(_dp),
                                                   self.frac.ctypes.data_as(_dp), kap.ctypes.data_as(_dp),
                                                   ctypes.byref(self.ieruption), self.jeruption.value,
                                                   flux.ctypes.data_as(_dp)))
        return flux

    def step(self, T, time: float, dt: float, ivent: int, jvent: int):
        """The source part of one time step (src/Ash3d.F90:406-431) on a
        Transport context or the oracle: returns the inserted flux or None."""
        if not self.pulses(time, dt):
            return None
        st = T.st
        kap = st.kappa_pd[ivent + 1, jvent + 1, 2:st.nz + 2]
        flux = self.nodes(kap, st.ns)
        T.source_column_add(ivent, jvent, flux)
        return flux


# ---- umbrella sources (src/Source_Umbrella.f90) -----------------------------

UmbrellaWindParams = _tr.UmbrellaWindParams


def _check(lib, rc):
    if rc:
        raise _tr.Ash3dError(rc, lib.ash3d_tr_last_error().decode())


def umbrella_stencil(kappa_pd, ivent: int, jvent: int):
    """Allocate_Source_Umbrella's weights: (avg_stencil (3,3,nz), scale_fac
    (nz)) from the padded kappa_pd (nx+4, ny+4, nz+4)."""
    lib = _tr.load()
    kap = np.asfortranarray(kappa_pd, dtype=np.float64)
    nx, ny, nz = (n - 4 for n in kap.shape)
    st = np.zeros((3, 3, nz), order="F")
    sf = np.zeros(nz)
    _check(lib, lib.ash3d_tr_umbrella_stencil(kap.ctypes.data_as(_dp), nx, ny, nz, ivent, jvent,
                                              st.ctypes.data_as(_dp), sf.ctypes.data_as(_dp)))
    return st, sf


def umbrella_levels(z_vec_init, plume_height: float):
    """(ibase, itop): the levels holding 0.75 and 1.0 of the plume height;
    z_vec_init = z(0:nz) in km."""
    lib = _tr.load()
    z = _arr(z_vec_init)
    ib, it = ctypes.c_int32(), ctypes.c_int32()
    _check(lib, lib.ash3d_tr_umbrella_levels(z.ctypes.data_as(_dp), len(z) - 1, float(plume_height),
                                             ctypes.byref(ib), ctypes.byref(it)))
    return ib.value, it.value


def umbrella_winds(params: UmbrellaWindParams, lon_cc, lat_cc):
    """The radial umbrella-cloud winds (uvx, uvy), (nx, ny) in km/hr."""
    lib = _tr.load()
    lon, lat = _arr(lon_cc), _arr(lat_cc)
    nx, ny = len(lon), len(lat)
    ux, uy = np.zeros((nx, ny), order="F"), np.zeros((nx, ny), order="F")
    _check(lib, lib.ash3d_tr_umbrella_winds(ctypes.byref(params), nx, ny, lon.ctypes.data_as(_dp),
                                            lat.ctypes.data_as(_dp), ux.ctypes.data_as(_dp),
                                            uy.ctypes.data_as(_dp)))
    return ux, uy


def umbrella_nodes(flux, ibase: int, itop: int, scale_fac, n_gs_max: int | None = None):
    """TephraSourceNodes_Umbrella: the 3x3 patch flux (3,3,nz+1,ns) from the
    column flux (nz+1, ns)."""
    lib = _tr.load()
    f = np.asfortranarray(flux, dtype=np.float64)
    nz1, ns = f.shape
    sf = _arr(scale_fac)
    fu = np.zeros((3, 3, nz1, ns), order="F")
    _check(lib, lib.ash3d_tr_umbrella_nodes(f.ctypes.data_as(_dp), nz1 - 1, ns, ns if n_gs_max is None else n_gs_max,
                                            ibase, itop, sf.ctypes.data_as(_dp), fu.ctypes.data_as(_dp)))
    return fu


def umbrella_vol_inc(flux, ibase: int, itop: int, avg_stencil, kappa_vent_col, dt: float, magma_density: float,
                     n_gs_max: int | None = None) -> float:
    """SourceVolInc_Umbrella: km3 DRE inserted in one step."""
    lib = _tr.load()
    f = np.asfortranarray(flux, dtype=np.float64)
    nz1, ns = f.shape
    st = np.asfortranarray(avg_stencil, dtype=np.float64)
    kap = _arr(kappa_vent_col)
    vol = ctypes.c_double()
    _check(lib, lib.ash3d_tr_umbrella_vol_inc(f.ctypes.data_as(_dp), nz1 - 1, ns, ns if n_gs_max is None else n_gs_max,
                                              ibase, itop, st.ctypes.data_as(_dp), kap.ctypes.data_as(_dp),
                                              float(dt), float(magma_density), ctypes.byref(vol)))
    return vol.value
