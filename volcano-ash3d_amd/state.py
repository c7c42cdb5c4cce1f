"""Host-side state of the Ash3d transport step, in the reference's layouts.

Every array is a Fortran-ordered float64 numpy array with the shape of the
reference module array it stands for, so its raw buffer is exactly what the
Fortran driver hands across the C ABI (src/Ash3d_VariableModules.f90:918-946,
src/Diffusion.F90:95-101, src/Topography.f90:412-413):

===================  ================================================  =====
array                reference declaration                             shape
===================  ================================================  =====
concen_pd            concen_pd(-1:nx+2,-1:ny+2,-1:nz+2,1:ns,ts0:ts1)   P,ns,2
vx/vy/vz/vh_pd       (-1:nx+2,-1:ny+2,-1:nz+2)                         P
vf_pd                (-1:nx+2,-1:ny+2,-1:nz+2,1:ns)                    P,ns
kappa_pd, sigma_n*   (-1:nx+2,-1:ny+2,-1:nz+2)                         P
kx, ky, kz           (0:nx+1,0:ny+1,0:nz+1)                            K
deldxond/deldyond    DelD{x,y}onD_cc(-1:nx+2,-1:ny+2)                  2-D
outflow_xz{1,2}_pd   (-1:nx+2,-1:nz+2,1:ns)
outflow_yz{1,2}_pd   (-1:ny+2,-1:nz+2,1:ns)
outflow_xy{1,2}_pd   (-1:nx+2,-1:ny+2,1:ns)
deposit_granularity  DepositGranularity(1:nx,1:ny,1:ns)
===================  ================================================  =====

(P = (nx+4, ny+4, nz+4), K = (nx+2, ny+2, nz+2).)  ``st.pad(i)`` maps a
reference index with lower bound -1 to the numpy index (i + 1).
"""
from __future__ import annotations

import dataclasses

import numpy as np

# Limiter ids; the reference selects one at compile time with -DLIM_*
# (src/makefile:126,169-190; src/AdvectionHorz_DCU.f90:178-226).
LIM_NONE, LIM_LAXWEN, LIM_BW, LIM_FROMM, LIM_MINMOD, LIM_SUPERBEE, LIM_MC = range(7)
LIMITER_NAMES = {
    "none": LIM_NONE, "laxwen": LIM_LAXWEN, "bw": LIM_BW, "fromm": LIM_FROMM,
    "minmod": LIM_MINMOD, "superbee": LIM_SUPERBEE, "mc": LIM_MC,
}

TS0, TS1 = 0, 1

_PAD_ARRAYS = ("vx_pd", "vy_pd", "vz_pd", "vh_pd", "kappa_pd",
               "sigma_nx_pd", "sigma_ny_pd", "sigma_nz_pd")
_ALL_ARRAYS = ("concen_pd",) + _PAD_ARRAYS + (
    "vf_pd", "kx", "ky", "kz", "deldxond_cc", "deldyond_cc",
    "outflow_xz1_pd", "outflow_xz2_pd", "outflow_yz1_pd", "outflow_yz2_pd",
    "outflow_xy1_pd", "outflow_xy2_pd", "deposit_granularity")


@dataclasses.dataclass
class TransportState:
    """All inputs and outputs of the hot path for one grid (one member)."""

    nx: int
    ny: int
    nz: int
    ns: int
    is_periodic: bool = False
    zscaling_id: int = 0
    limiter: int = LIM_SUPERBEE
    use_cn: bool = True
    dt: float = 0.0
    imp_fac: float = 0.5
    bounds: tuple = None  # (imin, imax, jmin, jmax, kmin, kmax)
    is_aloft: np.ndarray = None

    def __post_init__(self):
        nx, ny, nz, ns = self.nx, self.ny, self.nz, self.ns
        P = (nx + 4, ny + 4, nz + 4)
        f = dict(order="F", dtype=np.float64)
        self.concen_pd = np.zeros(P + (ns, 2), **f)
        for name in _PAD_ARRAYS:
            setattr(self, name, np.zeros(P, **f))
        # the reference allocates geometry initialised to 1.0 (Allocate_mesh)
        for name in ("kappa_pd", "sigma_nx_pd", "sigma_ny_pd", "sigma_nz_pd"):
            getattr(self, name)[...] = 1.0
        self.vf_pd = np.zeros(P + (ns,), **f)
        K = (nx + 2, ny + 2, nz + 2)
        self.kx = np.zeros(K, **f)
        self.ky = np.zeros(K, **f)
        self.kz = np.zeros(K, **f)
        self.deldxond_cc = np.zeros((nx + 4, ny + 4), **f)
        self.deldyond_cc = np.zeros((nx + 4, ny + 4), **f)
        self.outflow_xz1_pd = np.zeros((nx + 4, nz + 4, ns), **f)
        self.outflow_xz2_pd = np.zeros((nx + 4, nz + 4, ns), **f)
        self.outflow_yz1_pd = np.zeros((ny + 4, nz + 4, ns), **f)
        self.outflow_yz2_pd = np.zeros((ny + 4, nz + 4, ns), **f)
        self.outflow_xy1_pd = np.zeros((nx + 4, ny + 4, ns), **f)
        self.outflow_xy2_pd = np.zeros((nx + 4, ny + 4, ns), **f)
        self.deposit_granularity = np.zeros((nx, ny, ns), **f)
        if self.bounds is None:
            self.bounds = (1, nx, 1, ny, 1, nz)
        if self.is_aloft is None:
            self.is_aloft = np.ones(ns, dtype=np.int32)
        self.is_aloft = np.ascontiguousarray(self.is_aloft, dtype=np.int32)

    # -- helpers ---------------------------------------------------------
    @staticmethod
    def pad(i: int) -> int:
        """numpy index of reference index i of a (-1:...) dimension."""
        return i + 1

    def interior(self, t: int = TS0) -> np.ndarray:
        """View of concen_pd(1:nx,1:ny,1:nz,:,t)."""
        return self.concen_pd[2:self.nx + 2, 2:self.ny + 2, 2:self.nz + 2, :, t]

    def array_names(self):
        return _ALL_ARRAYS

    def copy(self) -> "TransportState":
        other = TransportState(self.nx, self.ny, self.nz, self.ns, self.is_periodic,
                               self.zscaling_id, self.limiter, self.use_cn, self.dt,
                               self.imp_fac, tuple(self.bounds), self.is_aloft.copy())
        for name in _ALL_ARRAYS:
            getattr(other, name)[...] = getattr(self, name)
        return other

    def arrays(self) -> dict:
        return {name: getattr(self, name) for name in _ALL_ARRAYS}

    def nbytes(self) -> int:
        return sum(a.nbytes for a in self.arrays().values())
