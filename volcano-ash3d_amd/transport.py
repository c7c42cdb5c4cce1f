"""ctypes mirror of the C ABI in include/ash3d_tr.h.

``Transport`` plays the role of the reference's transport modules for one
``TransportState``: it owns a device context (created from the state's
host arrays), and its methods carry the reference's routine names and
argument meaning -- ``Set_BC(bc_code)``, ``advect_x()``, ``advect_y()``,
``AdvectHorz(itoggle)``, ``AdvectVert()``, ``diff_x/y/z()``,
``diffCN_x/y/z()``, ``DiffuseHorz(i)``, ``DiffuseVert()`` -- reading dt,
IsAloft, Imp_fac and the imin..kmax bounds from the state the way the
Fortran routines read them from the modules.  Results stay on the device
until ``sync_to_host()``.

Errors from the library raise ``Ash3dError``; a missing library raises at
load time (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libash3d_tr.so")
ABI_VERSION = 1

SYNC_CONCEN, SYNC_WINDS, SYNC_VF, SYNC_GEOMETRY = 1, 2, 4, 8
SYNC_DIFFK, SYNC_TOPO, SYNC_OUTFLOW, SYNC_DEPOSIT = 16, 32, 64, 128
SYNC_ALL = 0xFF

ERRORS = {1: "EINVAL", 2: "EDEVICE", 3: "ENOMEM", 4: "EBC", 5: "ESINGULAR", 6: "ENOLIB"}


class Ash3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ash3d_tr error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Desc(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("ns", ctypes.c_int32),
        ("is_periodic", ctypes.c_int32), ("zscaling_id", ctypes.c_int32), ("limiter", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("concen_pd", _dp),
        ("vx_pd", _dp), ("vy_pd", _dp), ("vz_pd", _dp), ("vh_pd", _dp), ("vf_pd", _dp),
        ("kappa_pd", _dp), ("sigma_nx_pd", _dp), ("sigma_ny_pd", _dp), ("sigma_nz_pd", _dp),
        ("kx", _dp), ("ky", _dp), ("kz", _dp),
        ("deldxond_cc", _dp), ("deldyond_cc", _dp),
        ("outflow_xz1_pd", _dp), ("outflow_xz2_pd", _dp),
        ("outflow_yz1_pd", _dp), ("outflow_yz2_pd", _dp),
        ("outflow_xy1_pd", _dp), ("outflow_xy2_pd", _dp),
        ("deposit_granularity", _dp),
    ]


class DtParams(ctypes.Structure):
    """ash3d_tr_dt_params: the inputs of Adjust_DT (src/Adjust_dt.f90:89-333)."""
    _fields_ = [("is_latlon", ctypes.c_int32), ("use_met_next", ctypes.c_int32), ("n_gs_aloft", ctypes.c_int32),
                ("use_diffusion", ctypes.c_int32), ("use_cn", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.POINTER(ctypes.c_double)),
                ("imp_dt_fac", ctypes.c_double), ("cfl", ctypes.c_double), ("dt_min", ctypes.c_double),
                ("dt_max", ctypes.c_double), ("eps_small", ctypes.c_double), ("time", ctypes.c_double),
                ("next_write_time", ctypes.c_double), ("simtime", ctypes.c_double)]


class ProductParams(ctypes.Structure):
    """ash3d_tr_product_params: thresholds of ConcentrationCalculator and
    AshThicknessCalculator and the deposit density."""
    _fields_ = [("cloudcon_thresh", ctypes.c_double), ("cloudload_thresh", ctypes.c_double),
                ("depo_thresh", ctypes.c_double), ("deposit_density", ctypes.c_double)]


class ProductsOut(ctypes.Structure):
    """ash3d_tr_products_out."""
    _fields_ = [("cloud_load", ctypes.POINTER(ctypes.c_double)),
                ("max_concentration", ctypes.POINTER(ctypes.c_double)),
                ("max_height", ctypes.POINTER(ctypes.c_double)), ("min_height", ctypes.POINTER(ctypes.c_double)),
                ("deposit_thickness", ctypes.POINTER(ctypes.c_double)),
                ("mask_cloud", ctypes.POINTER(ctypes.c_int32)), ("mask_deposit", ctypes.POINTER(ctypes.c_int32)),
                ("cloud_area", ctypes.c_double), ("deposit_area", ctypes.c_double)]


class UmbrellaWindParams(ctypes.Structure):
    """ash3d_tr_umbrella_wind_params: the inputs of umbrella_winds
    (src/Source_Umbrella.f90:206-394)."""
    _fields_ = [("vel_model", ctypes.c_int32), ("k_entrainment", ctypes.c_double), ("lambda_", ctypes.c_double),
                ("n_bv", ctypes.c_double), ("mass_flux_rate", ctypes.c_double),
                ("airborne_one_bin", ctypes.c_int32), ("lat_volcano", ctypes.c_double),
                ("lon_volcano", ctypes.c_double), ("time", ctypes.c_double), ("dt", ctypes.c_double),
                ("first_time", ctypes.c_int32), ("simtime", ctypes.c_double), ("end_time", ctypes.c_double),
                ("de_km", ctypes.c_double), ("dn_km", ctypes.c_double), ("ivent", ctypes.c_int32),
                ("jvent", ctypes.c_int32), ("is_latlon", ctypes.c_int32)]


class SourceDesc(ctypes.Structure):
    """ash3d_tr_source_desc: eruptive pulses and vertical grid of the source
    builders (src/Source.f90)."""
    _fields_ = [("source_type", ctypes.c_int32), ("neruptions", ctypes.c_int32),
                ("start_time", ctypes.POINTER(ctypes.c_double)), ("duration", ctypes.POINTER(ctypes.c_double)),
                ("volume", ctypes.POINTER(ctypes.c_double)), ("plume_height", ctypes.POINTER(ctypes.c_double)),
                ("suzuki_a", ctypes.c_double), ("magma_density", ctypes.c_double), ("z_volcano", ctypes.c_double),
                ("zscaling_id", ctypes.c_int32), ("zsurf_vent", ctypes.c_double), ("ztop", ctypes.c_double),
                ("s_lb", ctypes.POINTER(ctypes.c_double)), ("s_cc", ctypes.POINTER(ctypes.c_double)),
                ("ds", ctypes.POINTER(ctypes.c_double)), ("bin_mass", ctypes.POINTER(ctypes.c_double)),
                ("prof_dz", ctypes.POINTER(ctypes.c_double)), ("prof_nz", ctypes.POINTER(ctypes.c_int32)),
                ("prof_volume", ctypes.POINTER(ctypes.c_double)), ("prof_ld", ctypes.c_int32)]


_fp = ctypes.POINTER(ctypes.c_float)

# every exported entry point and its argument types
_SIGS = {
    "ash3d_tr_create": [ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_void_p)],
    "ash3d_tr_destroy": [ctypes.c_void_p],
    "ash3d_tr_sync_from_host": [ctypes.c_void_p, ctypes.c_uint32],
    "ash3d_tr_sync_to_host": [ctypes.c_void_p, ctypes.c_uint32],
    "ash3d_tr_synchronize": [ctypes.c_void_p],
    "ash3d_tr_set_bc": [ctypes.c_void_p, ctypes.c_int32],
    "ash3d_tr_advect_x": [ctypes.c_void_p, ctypes.c_double, _ip, _ip],
    "ash3d_tr_advect_y": [ctypes.c_void_p, ctypes.c_double, _ip, _ip],
    "ash3d_tr_advect_z": [ctypes.c_void_p, ctypes.c_double, _ip, _ip],
    "ash3d_tr_advect_horz": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _ip],
    "ash3d_tr_advect_vert": [ctypes.c_void_p, ctypes.c_double, _ip],
    "ash3d_tr_diffuse_dir": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                             ctypes.c_double, _ip, _ip],
    "ash3d_tr_diffuse_horz": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                              ctypes.c_double, _ip],
    "ash3d_tr_diffuse_vert": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _ip],
    "ash3d_tr_step": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                      ctypes.c_double, _ip],
    "ash3d_tr_step_keep": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                           ctypes.c_double, _ip, ctypes.c_int32, ctypes.c_int32],
    "ash3d_tr_step_glue": [ctypes.c_void_p],
    "ash3d_tr_source_column_add": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _dp],
    "ash3d_tr_mass_totals": [ctypes.c_void_p, _dp, _dp, _dp],
    "ash3d_tr_mass_totals_noalloc": [ctypes.c_void_p, ctypes.c_void_p],
    "ash3d_tr_tune": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    "ash3d_tr_tune_get": [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)],
    "ash3d_tr_profile": [ctypes.c_void_p, ctypes.c_int32],
    "ash3d_tr_profile_read": [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_int64)],
    "ash3d_tr_checkpoints": [ctypes.c_void_p, ctypes.c_int32],
    "ash3d_tr_checkpoint": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    "ash3d_tr_adj_init": [ctypes.c_void_p, ctypes.c_int32],
    "ash3d_tr_adj_set": [ctypes.c_void_p, _dp],
    "ash3d_tr_adj_get": [ctypes.c_void_p, _dp],
    "ash3d_tr_adj_advect": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, _ip, _ip],
    "ash3d_tr_adj_diffuse": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                             ctypes.c_double, _ip, _ip],
    "ash3d_tr_adj_step": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                          ctypes.c_double, _ip],
    "ash3d_tr_adj_step_kept": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_double, _ip, ctypes.c_int32, ctypes.c_int32],
    "ash3d_tr_adj_source": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32],
    "ash3d_tr_adj_source_read": [ctypes.c_void_p, _dp],
    "ash3d_tr_column_load": [ctypes.c_void_p, _dp, _dp],
    "ash3d_tr_adj_column_load": [ctypes.c_void_p, _dp, _dp],
    "ash3d_tr_products": [ctypes.c_void_p, _dp, _dp, _ip, ctypes.POINTER(ProductParams), ctypes.POINTER(ProductsOut)],
    "ash3d_tr_write_esri_ascii": [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, _dp, _ip, ctypes.c_char_p,
                                  ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double],
    "ash3d_tr_esri_ascii_name": [ctypes.c_char_p, ctypes.c_double, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32],
    "ash3d_tr_source_column": [ctypes.POINTER(SourceDesc), ctypes.c_int32, _dp, _dp, _dp],
    "ash3d_tr_eruptive_pulses": [ctypes.POINTER(SourceDesc), _dp, ctypes.c_double, ctypes.c_double, _ip, _ip, _dp,
                                 _ip],
    "ash3d_tr_source_nodes": [ctypes.POINTER(SourceDesc), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                              _dp, _dp, _ip, ctypes.c_int32, _dp],
    "ash3d_tr_umbrella_stencil": [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, _dp, _dp],
    "ash3d_tr_umbrella_levels": [_dp, ctypes.c_int32, ctypes.c_double, _ip, _ip],
    "ash3d_tr_umbrella_winds": [ctypes.POINTER(UmbrellaWindParams), ctypes.c_int32, ctypes.c_int32, _dp, _dp, _dp,
                                _dp],
    "ash3d_tr_umbrella_nodes": [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_int32, _dp, _dp],
    "ash3d_tr_umbrella_vol_inc": [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_int32, _dp, _dp, ctypes.c_double, ctypes.c_double, _dp],
    "ash3d_tr_umbrella_source_add": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                     ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp],
    "ash3d_tr_umbrella_winds_add": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _dp, _dp],
    "ash3d_tr_graph_capture": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_double, _ip],
    "ash3d_tr_graph_launch": [ctypes.c_void_p, ctypes.c_int32],
    "ash3d_tr_ensemble_create": [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, _dp, ctypes.c_double,
                                 ctypes.POINTER(_ip), ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)],
    "ash3d_tr_ensemble_step": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32],
    "ash3d_tr_ensemble_synchronize": [ctypes.c_void_p],
    "ash3d_tr_ensemble_info": [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_int32)],
    "ash3d_tr_ensemble_destroy": [ctypes.c_void_p],
    "ash3d_tr_met_push": [ctypes.c_void_p, _fp, _fp, _fp, _fp],
    "ash3d_tr_met_interp": [ctypes.c_void_p, ctypes.c_double, ctypes.c_int32],
    "ash3d_tr_adjust_dt": [ctypes.c_void_p, ctypes.POINTER(DtParams), _dp, _dp],
    "ash3d_tr_minmax_index": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _ip],
    "ash3d_tr_step_subgrid": [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                              ctypes.c_double, _ip, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                              ctypes.c_double, _ip],
}
PROF_KINDS = ("advect_x", "advect_y", "advect_z", "diffuse_x", "diffuse_y", "diffuse_z", "bc", "glue",
              "transpose", "adj_advect", "adj_diffuse", "copy", "advect_hz")
CKPT_SAVE, CKPT_LOAD = 0, 1
EXPORTS = tuple(_SIGS) + ("ash3d_tr_last_error", "ash3d_tr_build_info")

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is missing: build it with `python volcano-ash3d_amd/build.py` "
                "(there is no CPU fallback)")
        lib = ctypes.CDLL(path)
        for name, args in _SIGS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.ash3d_tr_last_error.restype = ctypes.c_char_p
        lib.ash3d_tr_build_info.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(_dp)


class Transport:
    """Device context for one TransportState, with the reference's API."""

    def __init__(self, st, device: int = 0):
        self.lib = load()
        self.st = st
        d = Desc()
        d.abi_version = ABI_VERSION
        d.nx, d.ny, d.nz, d.ns = st.nx, st.ny, st.nz, st.ns
        d.is_periodic = int(bool(st.is_periodic))
        d.zscaling_id = int(st.zscaling_id)
        d.limiter = int(st.limiter)
        d.device = device
        for name in st.array_names():
            setattr(d, name, _ptr(getattr(st, name)))
        self._desc = d
        self._ctx = ctypes.c_void_p()
        self._check(self.lib.ash3d_tr_create(ctypes.byref(d), ctypes.byref(self._ctx)))

    # -- plumbing ----------------------------------------------------------
    def _check(self, rc):
        if rc != 0:
            raise Ash3dError(rc, self.lib.ash3d_tr_last_error().decode())

    def _aloft(self):
        self.st.is_aloft = np.ascontiguousarray(self.st.is_aloft, dtype=np.int32)
        return self.st.is_aloft.ctypes.data_as(_ip)

    def _bounds(self):
        self._b = np.ascontiguousarray(self.st.bounds, dtype=np.int32)
        return self._b.ctypes.data_as(_ip)

    def _full_bounds(self):
        st = self.st
        st.bounds = (1, st.nx, 1, st.ny, 1, st.nz)

    def close(self):
        if self._ctx:
            self.lib.ash3d_tr_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync_to_host(self, mask: int = SYNC_ALL):
        self._check(self.lib.ash3d_tr_sync_to_host(self._ctx, mask))

    def sync_from_host(self, mask: int = SYNC_ALL):
        self._check(self.lib.ash3d_tr_sync_from_host(self._ctx, mask))

    def synchronize(self):
        self._check(self.lib.ash3d_tr_synchronize(self._ctx))

    # -- the reference routines -------------------------------------------
    def Set_BC(self, bc_code: int):
        self._check(self.lib.ash3d_tr_set_bc(self._ctx, bc_code))

    def advect_x(self):
        self._check(self.lib.ash3d_tr_advect_x(self._ctx, self.st.dt, self._aloft(), self._bounds()))

    def advect_y(self):
        self._check(self.lib.ash3d_tr_advect_y(self._ctx, self.st.dt, self._aloft(), self._bounds()))

    def advect_z(self):
        self._check(self.lib.ash3d_tr_advect_z(self._ctx, self.st.dt, self._aloft(), self._bounds()))

    def AdvectHorz(self, itoggle: int):
        self._full_bounds()
        self._check(self.lib.ash3d_tr_advect_horz(self._ctx, itoggle, self.st.dt, self._aloft()))

    def AdvectVert(self):
        self._check(self.lib.ash3d_tr_advect_vert(self._ctx, self.st.dt, self._aloft()))

    def _diffuse(self, dirn: int, use_cn: bool):
        self._check(self.lib.ash3d_tr_diffuse_dir(self._ctx, dirn, int(use_cn), self.st.dt,
                                                  self.st.imp_fac, self._aloft(), self._bounds()))

    def diff_x(self):
        self._diffuse(0, False)

    def diff_y(self):
        self._diffuse(1, False)

    def diff_z(self):
        self._diffuse(2, False)

    def diffCN_x(self):
        self._diffuse(0, True)

    def diffCN_y(self):
        self._diffuse(1, True)

    def diffCN_z(self):
        self._diffuse(2, True)

    def DiffuseHorz(self, itoggle: int):
        self._full_bounds()
        self._check(self.lib.ash3d_tr_diffuse_horz(self._ctx, itoggle, int(bool(self.st.use_cn)), self.st.dt,
                                                   self.st.imp_fac, self._aloft()))

    def DiffuseVert(self):
        self._check(self.lib.ash3d_tr_diffuse_vert(self._ctx, int(bool(self.st.use_cn)), self.st.dt,
                                                   self.st.imp_fac, self._aloft()))

    # -- driver-level entry points ------------------------------------------
    def transport_step(self, itime: int, use_diffusion: bool):
        """Set_BC(1); AdvectHorz(itime); AdvectVert; [Set_BC(2);
        DiffuseHorz(itime); DiffuseVert] (src/Ash3d.F90:491-510)."""
        self._full_bounds()
        self._check(self.lib.ash3d_tr_step(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                           int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft()))

    def transport_step_keep(self, itime: int, use_diffusion: bool, slot_b: int, slot_z: int):
        """transport_step, keeping the inputs of the second horizontal and
        the vertical advection sweep in checkpoint slots slot_b, slot_z."""
        self._full_bounds()
        self._check(self.lib.ash3d_tr_step_keep(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                                int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft(),
                                                int(slot_b), int(slot_z)))

    def step_glue(self):
        self._check(self.lib.ash3d_tr_step_glue(self._ctx))

    def source_column_add(self, ivent: int, jvent: int, flux: np.ndarray):
        flux = np.asfortranarray(flux, dtype=np.float64)
        assert flux.shape == (self.st.nz + 1, self.st.ns)
        self._check(self.lib.ash3d_tr_source_column_add(self._ctx, ivent, jvent, self.st.dt,
                                                        flux.ctypes.data_as(_dp)))


    def umbrella_source_add(self, ivent: int, jvent: int, ibase: int, itop: int, n_gs_max: int,
                            avg_stencil: np.ndarray, flux_umb: np.ndarray, dt: float | None = None):
        """The umbrella source step of src/Ash3d.F90:447-460 on ts0:
        avg_stencil (3,3,nz), flux_umb (3,3,nz+1,ns) (see source.py)."""
        st = avg_stencil if avg_stencil.flags.f_contiguous else np.asfortranarray(avg_stencil)
        fu = np.asfortranarray(flux_umb, dtype=np.float64)
        assert st.shape == (3, 3, self.st.nz) and fu.shape == (3, 3, self.st.nz + 1, self.st.ns)
        self._check(self.lib.ash3d_tr_umbrella_source_add(self._ctx, ivent, jvent,
                                                          self.st.dt if dt is None else float(dt), ibase, itop,
                                                          n_gs_max, np.asfortranarray(st, dtype=np.float64)
                                                          .ctypes.data_as(_dp), fu.ctypes.data_as(_dp)))

    def umbrella_winds_add(self, ibase: int, itop: int, uvx: np.ndarray, uvy: np.ndarray):
        """vx_pd, vy_pd += the umbrella winds (nx, ny) on levels ibase..itop."""
        ux = np.asfortranarray(uvx, dtype=np.float64)
        uy = np.asfortranarray(uvy, dtype=np.float64)
        assert ux.shape == uy.shape == (self.st.nx, self.st.ny)
        self._check(self.lib.ash3d_tr_umbrella_winds_add(self._ctx, ibase, itop, ux.ctypes.data_as(_dp),
                                                         uy.ctypes.data_as(_dp)))

    def mass_totals(self):
        """Per-bin (aloft, deposit, outflow) masses in kg, computed on the
        device (src/Output_Vars.f90:1255-1459 before the volume conversion)."""
        ns = self.st.ns
        a, dpo, o = np.zeros(ns), np.zeros(ns), np.zeros(ns)
        self._check(self.lib.ash3d_tr_mass_totals(self._ctx, a.ctypes.data_as(_dp), dpo.ctypes.data_as(_dp),
                                                  o.ctypes.data_as(_dp)))
        return a, dpo, o

    def mass_totals_noalloc(self, out_ptr: int):
        """ash3d_tr_mass_totals_noalloc: the per-bin totals into device
        memory at out_ptr (3*ns doubles: aloft, deposit, outflow), queued on
        the context's stream; the library allocates nothing."""
        self._check(self.lib.ash3d_tr_mass_totals_noalloc(self._ctx, ctypes.c_void_p(int(out_ptr))))

    # -- captured steps ----------------------------------------------------
    def graph_capture(self, itime: int, use_diffusion: bool):
        """Capture step_glue + transport_step(itime) for itime's parity."""
        self._full_bounds()
        self._check(self.lib.ash3d_tr_graph_capture(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                                    int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft()))

    def graph_launch(self, itime: int):
        """step_glue + transport_step(itime) from the captured graph."""
        self._check(self.lib.ash3d_tr_graph_launch(self._ctx, itime))

    # -- checkpoints -------------------------------------------------------
    def checkpoints(self, nslots: int):
        """Resize the device checkpoint store to nslots copies of ts0."""
        self._check(self.lib.ash3d_tr_checkpoints(self._ctx, int(nslots)))

    def checkpoint_save(self, slot: int):
        self._check(self.lib.ash3d_tr_checkpoint(self._ctx, int(slot), CKPT_SAVE))

    def checkpoint_load(self, slot: int):
        self._check(self.lib.ash3d_tr_checkpoint(self._ctx, int(slot), CKPT_LOAD))

    # -- observation operator ----------------------------------------------
    def products(self, dz, zcc, bounds=None, cloudcon_thresh=1.0e-3, cloudload_thresh=0.2, depo_thresh=1.0e-2,
                 deposit_density=1000.0) -> dict:
        """ConcentrationCalculator + AshThicknessCalculator on the device
        state: a dict of (nx, ny) Fortran-order arrays (cloud_load,
        max_concentration, max_height, min_height, deposit_thickness,
        mask_cloud, mask_deposit) and the cloud / deposit areas."""
        nx, ny = self.st.nx, self.st.ny
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        zcc = np.ascontiguousarray(zcc, dtype=np.float64)
        assert dz.shape == (self.st.nz,) and zcc.shape == (self.st.nz,)
        res = {k: np.zeros((nx, ny), order="F") for k in ("cloud_load", "max_concentration", "max_height",
                                                            "min_height", "deposit_thickness")}
        res["mask_cloud"] = np.zeros((nx, ny), dtype=np.int32, order="F")
        res["mask_deposit"] = np.zeros((nx, ny), dtype=np.int32, order="F")
        o = ProductsOut()
        for k in ("cloud_load", "max_concentration", "max_height", "min_height", "deposit_thickness"):
            setattr(o, k, res[k].ctypes.data_as(_dp))
        o.mask_cloud = res["mask_cloud"].ctypes.data_as(_ip)
        o.mask_deposit = res["mask_deposit"].ctypes.data_as(_ip)
        prm = ProductParams(cloudcon_thresh, cloudload_thresh, depo_thresh, deposit_density)
        b = None if bounds is None else np.ascontiguousarray(bounds[:4], dtype=np.int32)
        self._check(self.lib.ash3d_tr_products(self._ctx, dz.ctypes.data_as(_dp), zcc.ctypes.data_as(_dp),
                                               None if b is None else b.ctypes.data_as(_ip), ctypes.byref(prm),
                                               ctypes.byref(o)))
        res["cloud_area"] = o.cloud_area
        res["deposit_area"] = o.deposit_area
        return res

    def column_load(self, dz: np.ndarray) -> np.ndarray:
        """Cloud load (t/km2) of every column, (nx, ny) Fortran order."""
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        assert dz.shape == (self.st.nz,)
        out = np.zeros((self.st.nx, self.st.ny), order="F")
        self._check(self.lib.ash3d_tr_column_load(self._ctx, dz.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
        return out

    # -- discrete adjoint --------------------------------------------------
    def adj_init(self, n_source_slots: int = 0):
        self._check(self.lib.ash3d_tr_adj_init(self._ctx, int(n_source_slots)))

    def adj_set(self, lam):
        """lam: array of concen_pd(:,:,:,:,ts0)'s shape, or None for zero."""
        if lam is None:
            self._check(self.lib.ash3d_tr_adj_set(self._ctx, None))
            return
        lam = np.asfortranarray(lam, dtype=np.float64)
        assert lam.shape == self.st.concen_pd.shape[:4]
        self._check(self.lib.ash3d_tr_adj_set(self._ctx, lam.ctypes.data_as(_dp)))

    def adj_get(self) -> np.ndarray:
        out = np.zeros(self.st.concen_pd.shape[:4], order="F")
        self._check(self.lib.ash3d_tr_adj_get(self._ctx, out.ctypes.data_as(_dp)))
        return out

    def adj_advect(self, dirn: int):
        """lambda <- J^T lambda for advect_x/y/z (dirn 0/1/2) at the state in ts0."""
        self._check(self.lib.ash3d_tr_adj_advect(self._ctx, int(dirn), self.st.dt, self._aloft(), self._bounds()))

    def adj_diffuse(self, dirn: int, use_cn: bool):
        self._check(self.lib.ash3d_tr_adj_diffuse(self._ctx, int(dirn), int(bool(use_cn)), self.st.dt,
                                                  self.st.imp_fac, self._aloft(), self._bounds()))

    def adj_step(self, itime: int, use_diffusion: bool):
        """Adjoint of transport_step(itime) with the step's input state in ts0."""
        self._full_bounds()
        self._check(self.lib.ash3d_tr_adj_step(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                               int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft()))

    def adj_step_kept(self, itime: int, use_diffusion: bool, slot_b: int, slot_z: int):
        """adj_step with the sweep inputs transport_step_keep kept."""
        self._full_bounds()
        self._check(self.lib.ash3d_tr_adj_step_kept(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                                    int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft(),
                                                    int(slot_b), int(slot_z)))

    def adj_source(self, ivent: int, jvent: int, slot: int):
        self._check(self.lib.ash3d_tr_adj_source(self._ctx, ivent, jvent, self.st.dt, int(slot)))

    def adj_source_read(self, n_slots: int) -> np.ndarray:
        """Source gradients, shape (n_slots, ns, nz+1)."""
        out = np.zeros((max(n_slots, 1), self.st.ns, self.st.nz + 1))
        self._check(self.lib.ash3d_tr_adj_source_read(self._ctx, out.ctypes.data_as(_dp)))
        return out[:n_slots]

    def adj_column_load(self, dz: np.ndarray, w: np.ndarray):
        """lambda += L^T w for the column load; w of shape (nx, ny)."""
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        w = np.asfortranarray(w, dtype=np.float64)
        assert w.shape == (self.st.nx, self.st.ny)
        self._check(self.lib.ash3d_tr_adj_column_load(self._ctx, dz.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))

    # -- driver work between the hot-path calls ------------------------------
    def met_push(self, vx, vy, vz, vf=None):
        """Upload the next met step (m/s, float32 Fortran arrays (nx,ny,nz); vf (nx,ny,nz,ns))."""
        st = self.st
        arrs = []
        for a, shape in ((vx, (st.nx, st.ny, st.nz)), (vy, (st.nx, st.ny, st.nz)), (vz, (st.nx, st.ny, st.nz)),
                         (vf, (st.nx, st.ny, st.nz, st.ns))):
            if a is None:
                arrs.append(None)
                continue
            a = np.asfortranarray(a, dtype=np.float32)
            assert a.shape == shape
            arrs.append(a)
        self._met_keep = arrs
        self._check(self.lib.ash3d_tr_met_push(self._ctx, *[None if a is None else a.ctypes.data_as(_fp)
                                                            for a in arrs]))

    def met_interp(self, frac: float, use_vf: bool):
        self._check(self.lib.ash3d_tr_met_interp(self._ctx, float(frac), int(bool(use_vf))))

    def adjust_dt(self, params: DtParams):
        """(dt, [vxmax_dx, vymax_dy, vzmax_dz, time_diffuse])."""
        dt = ctypes.c_double()
        diag = np.zeros(4)
        self._check(self.lib.ash3d_tr_adjust_dt(self._ctx, ctypes.byref(params), ctypes.byref(dt),
                                                diag.ctypes.data_as(_dp)))
        return dt.value, diag

    def minmax_index(self, ivent: int, jvent: int, kvent: int, thresh: float = 1e-7):
        b = np.zeros(6, dtype=np.int32)
        self._check(self.lib.ash3d_tr_minmax_index(self._ctx, ivent, jvent, kvent, float(thresh),
                                                   b.ctypes.data_as(_ip)))
        return tuple(int(v) for v in b)

    def transport_step_subgrid(self, itime: int, use_diffusion: bool, ivent: int, jvent: int, kvent: int,
                               thresh: float = 1e-7):
        """The step as built with FAST_SUBGRID; returns the advection box."""
        b = np.zeros(6, dtype=np.int32)
        self._check(self.lib.ash3d_tr_step_subgrid(self._ctx, itime, self.st.dt, int(bool(use_diffusion)),
                                                   int(bool(self.st.use_cn)), self.st.imp_fac, self._aloft(),
                                                   ivent, jvent, kvent, float(thresh), b.ctypes.data_as(_ip)))
        return tuple(int(v) for v in b)

    # -- launch tuning ----------------------------------------------------
    TUNE_PARAMS = ("nb_advect_x", "nb_advect_y", "nb_advect_z", "nb_diffuse_x", "nb_diffuse_y", "nb_diffuse_z",
                   "x_tile", "line_solver", "part_bins", "part_shape_x", "part_shape_y", "part_shape_z",
                   "adv_groups_y", "adv_groups_z", "adv_x_rows", "part_max_line",
                   "part_max_line_strided", "adv_fused", "jacobi_sweeps", "jacobi_last_y",
                   "jacobi_last_z", "jacobi_x", "jacobi_last_x")
    LINE_SOLVERS = ("exact", "partitioned")

    def tune(self, **kw):
        """Set launch parameters, e.g. tune(nb_advect_z=2); line_solver takes
        "exact" (dgtsv order, bitwise the reference) or "partitioned"."""
        for name, value in kw.items():
            if name == "line_solver" and isinstance(value, str):
                value = self.LINE_SOLVERS.index(value)
            self._check(self.lib.ash3d_tr_tune(self._ctx, self.TUNE_PARAMS.index(name), int(value)))

    def tune_get(self, name: str) -> int:
        """The current value of a launch setting (see tune)."""
        v = ctypes.c_int32()
        self._check(self.lib.ash3d_tr_tune_get(self._ctx, self.TUNE_PARAMS.index(name), ctypes.byref(v)))
        return int(v.value)

    def tuning(self) -> dict:
        """Every launch setting with its current value."""
        return {name: self.tune_get(name) for name in self.TUNE_PARAMS}

    # -- kernel timing ----------------------------------------------------
    def profile(self, enable: bool):
        self._check(self.lib.ash3d_tr_profile(self._ctx, int(bool(enable))))

    def profile_read(self) -> dict:
        out = {}
        for k, name in enumerate(PROF_KINDS):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            self._check(self.lib.ash3d_tr_profile_read(self._ctx, k, ctypes.byref(ms), ctypes.byref(n)))
            out[name] = (ms.value, n.value)
        return out


class Ensemble:
    """Batched steps of several Transport contexts on one device
    (ash3d_tr_ensemble_*): each step() does, for every member, what
    step_glue() + transport_step(itime, True) does, as one batched launch
    per sweep kind and kernel variant (one HIP graph with use_graph).
    Every member must have taken a step first, with the partitioned line
    solver and Crank-Nicolson diffusion; each member's dt, imp_fac and
    is_aloft are taken at construction; streams (0: the library's 4)
    stream groups of about equal work run side by side."""

    def __init__(self, members, imp_fac: float | None = None, streams: int = 0):
        self.lib = load()
        self.members = list(members)
        n = len(self.members)
        if imp_fac is None:
            imp_fac = float(self.members[0].st.imp_fac)
        ctxs = (ctypes.c_void_p * n)(*[m._ctx.value for m in self.members])
        self._dt = np.array([float(m.st.dt) for m in self.members], dtype=np.float64)
        self._masks = [np.ascontiguousarray(m.st.is_aloft, dtype=np.int32) for m in self.members]
        masks = (_ip * n)(*[a.ctypes.data_as(_ip) for a in self._masks])
        self._ens = ctypes.c_void_p()
        self._check(self.lib.ash3d_tr_ensemble_create(ctxs, n, self._dt.ctypes.data_as(_dp), float(imp_fac),
                                                      masks, int(streams), ctypes.byref(self._ens)))

    def _check(self, rc):
        if rc != 0:
            raise Ash3dError(rc, self.lib.ash3d_tr_last_error().decode())

    def step(self, itime: int, use_graph: bool = True):
        self._check(self.lib.ash3d_tr_ensemble_step(self._ens, int(itime), int(bool(use_graph))))

    def synchronize(self):
        self._check(self.lib.ash3d_tr_ensemble_synchronize(self._ens))

    def info(self, itime: int = 1):
        """(batched launches, the members' own launches) of one step."""
        b, m = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.ash3d_tr_ensemble_info(self._ens, int(itime), ctypes.byref(b), ctypes.byref(m)))
        return b.value, m.value

    def close(self):
        if self._ens:
            self.lib.ash3d_tr_ensemble_destroy(self._ens)
            self._ens = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def write_esri_ascii(path: str, field: np.ndarray, mask=None, fill_value: str = "-9999.", is_latlon=False,
                     xll=0.0, yll=0.0, dx=1.0, dy=1.0):
    """write_2D_ASCII of an (nx, ny) field (src/write_ASCII.f90:264-382)."""
    lib = load()
    f = np.asfortranarray(field, dtype=np.float64)
    nx, ny = f.shape
    m = None if mask is None else np.asfortranarray(mask, dtype=np.int32)
    rc = lib.ash3d_tr_write_esri_ascii(path.encode(), nx, ny, f.ctypes.data_as(_dp),
                                       None if m is None else m.ctypes.data_as(_ip), fill_value.encode(),
                                       int(bool(is_latlon)), float(xll), float(yll), float(dx), float(dy))
    if rc:
        raise Ash3dError(rc, lib.ash3d_tr_last_error().decode())


def esri_ascii_name(root: str, time_h: float = 0.0, is_final: bool = False) -> str:
    """The file name write_2D_ASCII gives a product root at an output time."""
    lib = load()
    buf = ctypes.create_string_buffer(256)
    rc = lib.ash3d_tr_esri_ascii_name(root.encode(), float(time_h), int(bool(is_final)), buf, 256)
    if rc:
        raise Ash3dError(rc, lib.ash3d_tr_last_error().decode())
    return buf.value.decode()


def build_info() -> str:
    return load().ash3d_tr_build_info().decode()
