"""ctypes wrapper of the C oracle (ash3d_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke entry point (as
the checker) and by bench.py's cpu_baseline leg.  The product path never
imports this module.

``Oracle`` mirrors the reference routine names: ``Set_BC``, ``advect_x``,
``advect_y``, ``advect_z``, ``AdvectHorz``, ``AdvectVert``, ``diff_x/y/z``,
``diffCN_x/y/z``, ``DiffuseHorz``, ``DiffuseVert`` and ``dgtsv``; each call
updates the arrays of a ``TransportState`` in place, exactly as the Fortran
routines update the module arrays.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "liboracle.so")
SOURCES = [os.path.join(HERE, "ash3d_oracle.c")]

# -ffp-contract=off: no fused multiply-add, so each operation rounds exactly
# as the reference's expression does.  OpenMP only splits independent rows.
CFLAGS = ["-O3", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
          "-Wall", "-Wextra"]


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = LIB_PATH + ".tmp"
    cmd = ["gcc", *CFLAGS, *SOURCES, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class _OrcState(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32), ("ns", ctypes.c_int32),
        ("is_periodic", ctypes.c_int32), ("zscaling_id", ctypes.c_int32),
        ("limiter", ctypes.c_int32), ("use_cn", ctypes.c_int32),
        ("imin", ctypes.c_int32), ("imax", ctypes.c_int32), ("jmin", ctypes.c_int32),
        ("jmax", ctypes.c_int32), ("kmin", ctypes.c_int32), ("kmax", ctypes.c_int32),
        ("dt", ctypes.c_double), ("imp_fac", ctypes.c_double),
        ("is_aloft", _ip),
        ("concen_pd", _dp),
        ("vx_pd", _dp), ("vy_pd", _dp), ("vz_pd", _dp), ("vh_pd", _dp), ("vf_pd", _dp),
        ("kappa_pd", _dp), ("sigma_nx_pd", _dp), ("sigma_ny_pd", _dp), ("sigma_nz_pd", _dp),
        ("kx", _dp), ("ky", _dp), ("kz", _dp),
        ("deldxond_cc", _dp), ("deldyond_cc", _dp),
        ("outflow_xz1_pd", _dp), ("outflow_xz2_pd", _dp),
        ("outflow_yz1_pd", _dp), ("outflow_yz2_pd", _dp),
        ("outflow_xy1_pd", _dp), ("outflow_xy2_pd", _dp),
        ("deposit_granularity", _dp),
        ("last_info", ctypes.c_int32),
    ]


_fp = ctypes.POINTER(ctypes.c_float)


class _MetStep(ctypes.Structure):
    _fields_ = [("vx", _fp), ("vy", _fp), ("vz", _fp), ("vf", _fp)]


class DtParams(ctypes.Structure):
    """orc_dt_params (Adjust_DT inputs); the same field order as
    ash3d_tr_dt_params of the library ABI."""
    _fields_ = [("is_latlon", ctypes.c_int32), ("use_met_next", ctypes.c_int32), ("n_gs_aloft", ctypes.c_int32),
                ("use_diffusion", ctypes.c_int32), ("use_cn", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", _dp),
                ("imp_dt_fac", ctypes.c_double), ("cfl", ctypes.c_double), ("dt_min", ctypes.c_double),
                ("dt_max", ctypes.c_double), ("eps_small", ctypes.c_double), ("time", ctypes.c_double),
                ("next_write_time", ctypes.c_double), ("simtime", ctypes.c_double)]


def _met(step):
    """(vx, vy, vz, vf) float32 Fortran arrays -> _MetStep (vf may be None)."""
    m = _MetStep()
    for name, a in zip(("vx", "vy", "vz", "vf"), step):
        if a is not None:
            assert a.dtype == np.float32 and a.flags.f_contiguous
            setattr(m, name, a.ctypes.data_as(_fp))
    return m


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(_OrcState)
        for name in ("orc_advect_x", "orc_advect_y", "orc_advect_z", "orc_advect_vert",
                     "orc_diff_x", "orc_diff_y", "orc_diff_z", "orc_diffcn_x", "orc_diffcn_y",
                     "orc_diffcn_z", "orc_diffuse_vert", "orc_step_glue"):
            getattr(lib, name).argtypes = [P]
            getattr(lib, name).restype = None
        for name in ("orc_advect_horz", "orc_diffuse_horz"):
            getattr(lib, name).argtypes = [P, ctypes.c_int]
            getattr(lib, name).restype = None
        lib.orc_set_bc.argtypes = [P, ctypes.c_int]
        lib.orc_set_bc.restype = ctypes.c_int
        lib.orc_transport_step.argtypes = [P, ctypes.c_int, ctypes.c_int]
        lib.orc_transport_step.restype = ctypes.c_int
        lib.orc_source_column_add.argtypes = [P, ctypes.c_int, ctypes.c_int, _dp]
        lib.orc_source_column_add.restype = None
        lib.orc_umbrella_source_add.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, _dp, _dp]
        lib.orc_umbrella_source_add.restype = None
        lib.orc_umbrella_winds_add.argtypes = [P, ctypes.c_int, ctypes.c_int, _dp, _dp]
        lib.orc_umbrella_winds_add.restype = None
        lib.orc_dgtsv.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int)]
        lib.orc_dgtsv.restype = None
        lib.orc_limiter_phi.argtypes = [ctypes.c_int, ctypes.c_double]
        lib.orc_limiter_phi.restype = ctypes.c_double
        lib.orc_num_threads.restype = ctypes.c_int
        lib.orc_set_num_threads.argtypes = [ctypes.c_int]
        lib.orc_set_num_threads.restype = None
        M = ctypes.POINTER(_MetStep)
        lib.orc_met_interp.argtypes = [P, M, M, ctypes.c_double, ctypes.c_int]
        lib.orc_met_interp.restype = None
        lib.orc_adjust_dt.argtypes = [P, ctypes.POINTER(DtParams), M, _dp]
        lib.orc_adjust_dt.restype = ctypes.c_double
        lib.orc_minmax_index.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _ip]
        lib.orc_minmax_index.restype = ctypes.c_int
        lib.orc_transport_step_subgrid.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   ctypes.c_int, ctypes.c_double, _ip]
        lib.orc_transport_step_subgrid.restype = ctypes.c_int
        lib.orc_products.argtypes = [P, _dp, _dp, _dp, _dp, _ip, _dp]
        lib.orc_products.restype = None
        _lib = lib
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


class Oracle:
    """The reference routines on a TransportState (host arrays, in place)."""

    def __init__(self, st):
        self.st = st
        self.lib = load()
        self._s = _OrcState()
        self._sync_in()

    def _sync_in(self):
        st, s = self.st, self._s
        s.nx, s.ny, s.nz, s.ns = st.nx, st.ny, st.nz, st.ns
        s.is_periodic = int(bool(st.is_periodic))
        s.zscaling_id = int(st.zscaling_id)
        s.limiter = int(st.limiter)
        s.use_cn = int(bool(st.use_cn))
        s.imin, s.imax, s.jmin, s.jmax, s.kmin, s.kmax = (int(b) for b in st.bounds)
        s.dt = float(st.dt)
        s.imp_fac = float(st.imp_fac)
        st.is_aloft = np.ascontiguousarray(st.is_aloft, dtype=np.int32)
        s.is_aloft = st.is_aloft.ctypes.data_as(_ip)
        for name in st.array_names():
            setattr(s, name, _ptr(getattr(st, name)))

    def _sync_out(self):
        s = self._s
        self.st.bounds = (s.imin, s.imax, s.jmin, s.jmax, s.kmin, s.kmax)

    def _call(self, name, *args):
        self._sync_in()
        r = getattr(self.lib, name)(ctypes.byref(self._s), *args)
        self._sync_out()
        return r

    # reference-named entry points
    def Set_BC(self, bc_code: int) -> int:
        return self._call("orc_set_bc", bc_code)

    def advect_x(self):
        self._call("orc_advect_x")

    def advect_y(self):
        self._call("orc_advect_y")

    def advect_z(self):
        self._call("orc_advect_z")

    def AdvectHorz(self, itoggle: int):
        self._call("orc_advect_horz", itoggle)

    def AdvectVert(self):
        self._call("orc_advect_vert")

    def diff_x(self):
        self._call("orc_diff_x")

    def diff_y(self):
        self._call("orc_diff_y")

    def diff_z(self):
        self._call("orc_diff_z")

    def diffCN_x(self):
        self._call("orc_diffcn_x")

    def diffCN_y(self):
        self._call("orc_diffcn_y")

    def diffCN_z(self):
        self._call("orc_diffcn_z")

    def DiffuseHorz(self, itoggle: int):
        self._call("orc_diffuse_horz", itoggle)

    def DiffuseVert(self):
        self._call("orc_diffuse_vert")

    def step_glue(self):
        self._call("orc_step_glue")

    def source_column_add(self, ivent: int, jvent: int, flux: np.ndarray):
        flux = np.asfortranarray(flux, dtype=np.float64)
        self._call("orc_source_column_add", ivent, jvent, _ptr(flux))

    def umbrella_source_add(self, ivent: int, jvent: int, ibase: int, itop: int, n_gs_max: int, avg_stencil,
                            flux_umb):
        st = np.asfortranarray(avg_stencil, dtype=np.float64)
        fu = np.asfortranarray(flux_umb, dtype=np.float64)
        self._call("orc_umbrella_source_add", ivent, jvent, ibase, itop, n_gs_max, _ptr(st), _ptr(fu))

    def umbrella_winds_add(self, ibase: int, itop: int, uvx, uvy):
        ux = np.asfortranarray(uvx, dtype=np.float64)
        uy = np.asfortranarray(uvy, dtype=np.float64)
        self._call("orc_umbrella_winds_add", ibase, itop, _ptr(ux), _ptr(uy))

    def transport_step(self, itime: int, use_diffusion: bool) -> int:
        return self._call("orc_transport_step", itime, int(bool(use_diffusion)))

    # driver work between the hot-path calls
    def met_interp(self, last, nxt, frac: float, use_vf: bool):
        """MesoInterpolater's wind time interpolation (src/MesoInterpolater.f90:203-230)."""
        a, b = _met(last), _met(nxt)
        self._call("orc_met_interp", ctypes.byref(a), ctypes.byref(b), float(frac), int(bool(use_vf)))

    def adjust_dt(self, params: DtParams, nxt=None):
        """Adjust_DT (src/Adjust_dt.f90:89-333): (dt, [vxmax_dx, vymax_dy, vzmax_dz, time_diffuse])."""
        m = _met(nxt) if nxt is not None else _MetStep()
        diag = np.zeros(4)
        dt = self._call("orc_adjust_dt", ctypes.byref(params), ctypes.byref(m), _ptr(diag))
        return dt, diag

    def minmax_index(self, ivent: int, jvent: int, kvent: int, thresh: float = 1e-7):
        """get_minmax_index (src/Calc_Mesh.f90:432-542): (status, bounds)."""
        b = np.zeros(6, dtype=np.int32)
        rc = self._call("orc_minmax_index", ivent, jvent, kvent, float(thresh), b.ctypes.data_as(_ip))
        return rc, tuple(int(v) for v in b)

    def transport_step_subgrid(self, itime: int, use_diffusion: bool, ivent: int, jvent: int, kvent: int,
                               thresh: float = 1e-7):
        b = np.zeros(6, dtype=np.int32)
        rc = self._call("orc_transport_step_subgrid", itime, int(bool(use_diffusion)), ivent, jvent, kvent,
                        float(thresh), b.ctypes.data_as(_ip))
        return rc, tuple(int(v) for v in b)

    def products(self, dz, zcc, cloudcon_thresh=1.0e-3, cloudload_thresh=0.2, depo_thresh=1.0e-2,
                 deposit_density=1000.0) -> dict:
        """ConcentrationCalculator + AshThicknessCalculator over the state's
        imin..imax x jmin..jmax columns (same keys as the library's)."""
        nx, ny = self.st.nx, self.st.ny
        dz = np.ascontiguousarray(dz, dtype=np.float64)
        zcc = np.ascontiguousarray(zcc, dtype=np.float64)
        prm = np.array([cloudcon_thresh, cloudload_thresh, depo_thresh, deposit_density])
        out = np.zeros(5 * nx * ny)
        masks = np.zeros(2 * nx * ny, dtype=np.int32)
        areas = np.zeros(2)
        self._call("orc_products", _ptr(dz), _ptr(zcc), _ptr(prm), _ptr(out), masks.ctypes.data_as(_ip), _ptr(areas))
        planes = out.reshape(5, ny, nx)
        res = {k: np.asfortranarray(planes[m].T) for m, k in enumerate(
            ("cloud_load", "max_concentration", "max_height", "min_height", "deposit_thickness"))}
        mk = masks.reshape(2, ny, nx)
        res["mask_cloud"] = np.asfortranarray(mk[0].T)
        res["mask_deposit"] = np.asfortranarray(mk[1].T)
        res["cloud_area"], res["deposit_area"] = float(areas[0]), float(areas[1])
        return res

    @property
    def last_info(self) -> int:
        return self._s.last_info


def dgtsv(dl, d, du, b):
    """LAPACK dgtsv restated (returns x, info and the factor arrays)."""
    lib = load()
    dl = np.array(dl, dtype=np.float64)
    d = np.array(d, dtype=np.float64)
    du = np.array(du, dtype=np.float64)
    b = np.array(b, dtype=np.float64, order="F")
    if b.ndim == 1:
        b = b.reshape(-1, 1, order="F")
    n, nrhs = b.shape
    info = ctypes.c_int(0)
    lib.orc_dgtsv(n, nrhs, _ptr(dl), _ptr(d), _ptr(du), _ptr(b), max(1, n), ctypes.byref(info))
    return b, info.value, dl, d, du


def limiter_phi(limiter: int, theta: float) -> float:
    return load().orc_limiter_phi(limiter, theta)


def num_threads() -> int:
    return load().orc_num_threads()


def set_num_threads(n: int) -> None:
    load().orc_set_num_threads(int(n))
