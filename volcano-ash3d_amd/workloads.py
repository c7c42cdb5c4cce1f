"""Synthetic inputs for the transport step: mesh geometry, winds, settling
velocities, diffusivities, time step and the standard workloads.

Geometry follows the reference's mesh set-up (src/Calc_Mesh.f90:78-168);
the time step follows Adjust_DT (src/Adjust_dt.f90:89-333, without the
output-time clipping).  Velocities are in km/h, diffusivities in km2/h, fall
velocities negative downward (src/MesoInterpolater.f90:205-206,
src/Input_Data.f90:3669).
"""
from __future__ import annotations

import math

import numpy as np

from .state import LIM_SUPERBEE, TransportState

# src/Ash3d_VariableModules.f90:276-292
PI = 3.141592653589793
DEG2RAD = 1.7453292519943295e-2
RAD_EARTH = 6371.229
MPS_2_KMPHR = 3.6
M2PS_2_KM2PHR = 3.6e-3
CFL = 0.8
DT_MIN = 1.0e-5
DT_MAX = 1.0


def z_levels(dz_vec):
    """dz_vec_pd, z_cc_pd on -1..nz+2 from interior layer thicknesses
    (src/Calc_Mesh.f90:78-101).  Returns arrays indexed k+1."""
    dz = np.asarray(dz_vec, dtype=np.float64)
    nz = dz.size
    z_init = np.concatenate([[0.0], np.cumsum(dz)])
    dzp = np.zeros(nz + 4)
    zcc = np.zeros(nz + 4)
    for k in range(1, nz + 1):
        dzp[k + 1] = z_init[k] - z_init[k - 1]
        zcc[k + 1] = z_init[k] - 0.5 * dzp[k + 1]
    zcc[0 + 1] = zcc[1 + 1] - dzp[1 + 1]
    zcc[-1 + 1] = zcc[0 + 1] - dzp[1 + 1]
    zcc[nz + 2] = zcc[nz + 1] + dzp[nz + 1]
    zcc[nz + 3] = zcc[nz + 1] + dzp[nz + 1] * 2.0
    dzp[0] = dzp[2]
    dzp[1] = dzp[2]
    dzp[nz + 2] = dzp[nz + 1]
    dzp[nz + 3] = dzp[nz + 1]
    return dzp, zcc


def cartesian_mesh(st: TransportState, dx: float, dy: float, dz_vec) -> None:
    """Projected-grid face areas and volumes (src/Calc_Mesh.f90:150-168)."""
    dzp, _ = z_levels(dz_vec)
    st.sigma_nz_pd[...] = dy * dx
    for kk in range(st.nz + 4):
        st.sigma_nx_pd[:, :, kk] = dzp[kk] * dy
        st.sigma_ny_pd[:, :, kk] = dzp[kk] * dx
        st.kappa_pd[:, :, kk] = dx * dy * dzp[kk]
    st.dx, st.dy = dx, dy
    st.dz_vec_pd = dzp
    st.is_latlon = False


def lonlat_mesh(st: TransportState, lonLL: float, latLL: float, de: float, dn: float,
                dz_vec) -> None:
    """Spherical-shell face areas and volumes incl. ghost rows
    (src/Calc_Mesh.f90:101-149).  Identical in i."""
    dzp, zcc = z_levels(dz_vec)
    ny, nz = st.ny, st.nz
    lat_cc = np.array([latLL + dn * float(j) - dn * 0.5 for j in range(-1, ny + 3)])
    for kk in range(nz + 4):
        r_1 = RAD_EARTH + zcc[kk] - 0.5 * dzp[kk]
        r_2 = RAD_EARTH + zcc[kk] + 0.5 * dzp[kk]
        rr_1 = r_1 * r_1
        drr = (r_2 * r_2 - r_1 * r_1)
        drrr = (r_2 * r_2 * r_2 - r_1 * r_1 * r_1)
        for jj in range(ny + 4):
            phi_1 = DEG2RAD * (lat_cc[jj] - 0.5 * dn)
            phi_2 = DEG2RAD * (lat_cc[jj] + 0.5 * dn)
            theta_1 = 0.5 * PI - phi_1
            theta_2 = 0.5 * PI - phi_2
            del_theta = dn * DEG2RAD
            del_costheta = math.cos(theta_2) - math.cos(theta_1)
            del_lam = de * DEG2RAD
            st.sigma_nx_pd[:, jj, kk] = 0.5 * del_theta * drr
            st.sigma_ny_pd[:, jj, kk] = 0.5 * math.sin(theta_1) * del_lam * drr
            st.sigma_nz_pd[:, jj, kk] = rr_1 * del_lam * del_costheta
            st.kappa_pd[:, jj, kk] = del_lam * drrr * del_costheta / 3.0
    phi_bot = DEG2RAD * (lat_cc[1 + 1] - dn * 0.5)
    phi_top = DEG2RAD * (lat_cc[ny + 1 + 1] + dn * 0.5)
    phi = max(abs(phi_bot), abs(phi_top))
    st.dx = RAD_EARTH * de * DEG2RAD * math.cos(phi)
    st.dy = RAD_EARTH * dn * DEG2RAD
    st.dz_vec_pd = dzp
    st.lat_cc = lat_cc
    st.z_cc = zcc
    st.is_latlon = True
    st.is_periodic = bool(de * st.nx >= 360.0 or abs(de * st.nx - 360.0) < 1e-12)


def correlated_noise(nx: int, ny: int, dx_km: float, dy_km: float, length_km: float,
                     sigma: float, rng: np.random.Generator) -> np.ndarray:
    """Seeded Gaussian-correlated 2-D field with standard deviation sigma and
    Gaussian correlation length length_km (spectral filter of white noise)."""
    w = rng.standard_normal((nx, ny))
    kx = np.fft.fftfreq(nx, d=dx_km) * 2 * np.pi
    ky = np.fft.fftfreq(ny, d=dy_km) * 2 * np.pi
    k2 = kx[:, None] ** 2 + ky[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-0.25 * k2 * length_km ** 2)))
    s = f.std()
    return f * (sigma / s) if s > 0 else f


def adjust_dt(st: TransportState, use_diffusion: bool, cfl: float = CFL,
              imp_dt_fac: float = 4.0) -> float:
    """CFL / diffusion-limited dt in hours (src/Adjust_dt.f90:128-323),
    rounded down to a multiple of DT_MIN."""
    nx, ny, nz = st.nx, st.ny, st.nz
    I = np.s_[2:nx + 2, 2:ny + 2, 2:nz + 2]
    kap = st.kappa_pd[I]
    dzp = st.dz_vec_pd[2:nz + 2]
    if not getattr(st, "is_latlon", False):
        vxmax_dx = np.abs(st.vx_pd[I]).max() / st.dx
        vymax_dy = np.abs(st.vy_pd[I]).max() / st.dy
    else:
        sx = np.minimum(st.sigma_nx_pd[2:nx + 2, 2:ny + 2, 2:nz + 2], st.sigma_nx_pd[3:nx + 3, 2:ny + 2, 2:nz + 2])
        sy = np.minimum(st.sigma_ny_pd[2:nx + 2, 2:ny + 2, 2:nz + 2], st.sigma_ny_pd[2:nx + 2, 3:ny + 3, 2:nz + 2])
        vxmax_dx = (np.abs(st.vx_pd[I]) * sx / kap).max()
        vymax_dy = (np.abs(st.vy_pd[I]) * sy / kap).max()
    # max over bins one bin at a time (no temporary of the whole vf array)
    vfmax = np.abs(st.vf_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, 0])
    for n in range(1, st.ns):
        np.maximum(vfmax, np.abs(st.vf_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, n]), out=vfmax)
    vzmax_dz = ((np.abs(st.vz_pd[I]) + vfmax) / dzp[None, None, :]).max()
    time_advect = 1.0 / max(vxmax_dx, vymax_dy, vzmax_dz)
    time_diffuse = DT_MAX
    if use_diffusion:
        def term(sig, kk, ax):
            sl = [slice(2, nx + 2), slice(2, ny + 2), slice(2, nz + 2)]
            sl2 = list(sl)
            sl2[ax] = slice(sl[ax].start + 1, sl[ax].stop + 1)
            ms = np.minimum(sig[tuple(sl)], sig[tuple(sl2)])
            return ((kap / ms) ** 2.0) / kk[1:nx + 1, 1:ny + 1, 1:nz + 1]
        with np.errstate(divide="ignore"):
            tmp4 = term(st.sigma_nx_pd, st.kx, 0) + term(st.sigma_ny_pd, st.ky, 1) + term(st.sigma_nz_pd, st.kz, 2)
        time_diffuse = min(DT_MAX, float(tmp4.min()))
        time_diffuse *= imp_dt_fac if st.use_cn else 0.5
    dt_tmp = min(time_advect, time_diffuse)
    if dt_tmp < DT_MIN:
        dt = DT_MIN
    elif dt_tmp > DT_MAX:
        dt = DT_MAX
    else:
        dt = min(1.0, cfl) * dt_tmp
    fac = int(dt / DT_MIN)
    return DT_MIN * fac


def fall_velocities(ns: int, vmin_mps: float, vmax_mps: float) -> np.ndarray:
    """Per-bin settling speeds, log-spaced, as negative (downward) km/h."""
    if ns == 1:
        v = np.array([vmin_mps])
    else:
        v = np.exp(np.linspace(math.log(vmin_mps), math.log(vmax_mps), ns))
    return -v * MPS_2_KMPHR


def set_uniform_fall(st: TransportState, vf_kmph) -> None:
    for n in range(st.ns):
        st.vf_pd[..., n] = vf_kmph[n]


def set_diffusivity(st: TransportState, kh_m2ps: float, kz_m2ps: float) -> None:
    """Allocate_Diff: kx=ky=Kh, kz=Kz (src/Diffusion.F90:99-101), km2/h."""
    st.kx[...] = kh_m2ps * M2PS_2_KM2PHR
    st.ky[...] = kh_m2ps * M2PS_2_KM2PHR
    st.kz[...] = kz_m2ps * M2PS_2_KM2PHR


def gaussian_blob(st: TransportState, center, widths, peak, bins=None) -> None:
    """A smooth positive cloud in ts0, used by tests as initial condition."""
    nx, ny, nz = st.nx, st.ny, st.nz
    i = np.arange(1, nx + 1)[:, None, None]
    j = np.arange(1, ny + 1)[None, :, None]
    k = np.arange(1, nz + 1)[None, None, :]
    g = np.exp(-0.5 * (((i - center[0]) / widths[0]) ** 2 + ((j - center[1]) / widths[1]) ** 2 +
                       ((k - center[2]) / widths[2]) ** 2))
    for n in (range(st.ns) if bins is None else bins):
        st.concen_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, n, 0] = peak * g * (1.0 + 0.1 * n)


# ---------------------------------------------------------------------------
# Standard workloads
# ---------------------------------------------------------------------------

def config1_state() -> TransportState:
    """Projected 100x100x30, dx=dy=5 km, dz=1 km, u=10, v=5, w=0 m/s, one bin
    with vf=-0.5 m/s, no diffusion, superbee, CFL 0.8."""
    st = TransportState(100, 100, 30, 1, limiter=LIM_SUPERBEE, use_cn=True)
    cartesian_mesh(st, 5.0, 5.0, np.ones(30))
    st.vx_pd[...] = 10.0 * MPS_2_KMPHR
    st.vy_pd[...] = 5.0 * MPS_2_KMPHR
    st.vz_pd[...] = 0.0
    set_uniform_fall(st, fall_velocities(1, 0.5, 0.5))
    st.dt = adjust_dt(st, use_diffusion=False)
    st.use_diffusion = False
    # point source: all mass in the cell containing 10 km at the vent (51,51)
    st.vent = (51, 51)
    st.source_k = 10
    st.source_rate = 1.0e6 * 3600.0  # kg/h for 1 h
    return st


def config1_source_flux(st: TransportState) -> np.ndarray:
    """SourceNodeFlux(1:nz+1,1:ns) in kg/km3/h for the config-1 point source."""
    flux = np.zeros((st.nz + 1, st.ns), order="F")
    i, j = st.vent
    flux[st.source_k - 1, :] = st.source_rate / st.kappa_pd[i + 1, j + 1, st.source_k + 1] / st.ns
    return flux


def zonal_jet_winds(st: TransportState, rng: np.random.Generator, jet_mps: float = 40.0,
                    sigma_mps: float = 3.0, length_km: float = 200.0) -> None:
    """u(z)=jet*sin(pi z/24 km) plus Gaussian-correlated perturbations on u
    and v (one field per level pair), w = 0."""
    nx, ny, nz = st.nx, st.ny, st.nz
    zc = st.z_cc[2:nz + 2]
    du = correlated_noise(nx, ny, st.dx, st.dy, length_km, sigma_mps, rng)
    dv = correlated_noise(nx, ny, st.dx, st.dy, length_km, sigma_mps, rng)
    jet = jet_mps * np.sin(np.pi * zc / 24.0)
    st.vx_pd[2:nx + 2, 2:ny + 2, 2:nz + 2] = (jet[None, None, :] + du[:, :, None]) * MPS_2_KMPHR
    st.vy_pd[2:nx + 2, 2:ny + 2, 2:nz + 2] = (dv[:, :, None] + 0.0 * zc[None, None, :]) * MPS_2_KMPHR
    st.vz_pd[...] = 0.0


def config2_state(seed: int = 2, ns: int = 15, nx: int = 360, ny: int = 240, nz: int = 50) -> TransportState:
    """Lon-lat 0.25 deg 360x240x50 (20-80 N), dz=0.5 km, 15 bins with |vf|
    log-spaced 0.01-5 m/s, zonal jet + seeded perturbations, Kh=Kz=500 m2/s,
    Crank-Nicolson."""
    st = TransportState(nx, ny, nz, ns, limiter=LIM_SUPERBEE, use_cn=True, imp_fac=0.5)
    lonlat_mesh(st, 180.0, 20.0, 0.25, 0.25, np.full(nz, 0.5))
    rng = np.random.default_rng(seed)
    zonal_jet_winds(st, rng)
    set_uniform_fall(st, fall_velocities(ns, 0.01, 5.0))
    set_diffusivity(st, 500.0, 500.0)
    st.use_diffusion = True
    st.dt = adjust_dt(st, use_diffusion=True)
    # an eruption cloud already aloft: Suzuki-like column spread by the winds
    gaussian_blob(st, (nx * 0.3, ny * 0.5, nz * 0.5), (nx * 0.08, ny * 0.1, nz * 0.15), 1.0e3)
    return st


def ensemble_member(member: int, ns: int = 10, size_scale: float = 1.0, seed: int = 3) -> TransportState:
    """One member of the ensemble workload: a lon-lat 0.25 deg grid whose
    size is drawn uniformly in [120,400] x [80,300] x [40,60] cells (times
    size_scale horizontally), centred at 20 N (every row within 57.5 S..57.5
    N), dz 0.5 km, ns bins with |vf| log-spaced
    0.02-3 m/s, the zonal jet with the member's own perturbations (sigma
    2 m/s), Kh = Kz = 1000 m2/s, Crank-Nicolson, and a cloud aloft around a
    plume height drawn in 8-20 km."""
    rng = np.random.default_rng([seed, member])
    nx = max(8, int(rng.integers(120, 401) * size_scale))
    ny = max(8, int(rng.integers(80, 301) * size_scale))
    nz = int(rng.integers(40, 61))
    st = TransportState(nx, ny, nz, ns, limiter=LIM_SUPERBEE, use_cn=True, imp_fac=0.5)
    lonlat_mesh(st, 180.0, 20.0 - 0.125 * ny, 0.25, 0.25, np.full(nz, 0.5))
    zonal_jet_winds(st, rng, sigma_mps=2.0)
    set_uniform_fall(st, fall_velocities(ns, 0.02, 3.0))
    set_diffusivity(st, 1000.0, 1000.0)
    st.use_diffusion = True
    st.dt = adjust_dt(st, use_diffusion=True)
    height_km = rng.uniform(8.0, 20.0)
    k = min(nz - 3.0, height_km / 0.5)
    gaussian_blob(st, (nx * 0.3, ny * 0.5, k), (nx * 0.08, ny * 0.1, nz * 0.12), 1.0e3)
    return st


def cloud_window(st: TransportState, center, widths, peak, cutoff: float = 6.0) -> None:
    """A Gaussian cloud in ts0 (bin n scaled by 1 + 0.1 n), written only
    inside +-cutoff widths of its centre (zero beyond), so that building
    it costs the window, not the grid."""
    nx, ny, nz = st.nx, st.ny, st.nz
    lo = [max(1, int(math.floor(c - cutoff * w))) for c, w in zip(center, widths)]
    hi = [min(m, int(math.ceil(c + cutoff * w))) for c, w, m in zip(center, widths, (nx, ny, nz))]
    i = np.arange(lo[0], hi[0] + 1)[:, None, None]
    j = np.arange(lo[1], hi[1] + 1)[None, :, None]
    k = np.arange(lo[2], hi[2] + 1)[None, None, :]
    g = np.exp(-0.5 * (((i - center[0]) / widths[0]) ** 2 + ((j - center[1]) / widths[1]) ** 2 +
                       ((k - center[2]) / widths[2]) ** 2))
    for n in range(st.ns):
        st.concen_pd[lo[0] + 1:hi[0] + 2, lo[1] + 1:hi[1] + 2, lo[2] + 1:hi[2] + 2, n, 0] = peak * g * (1.0 + 0.1 * n)


def global_jets_winds(st: TransportState, rng: np.random.Generator, jets=((40.0, 40.0), (30.0, -45.0)),
                      jet_width_deg: float = 12.0, sigma_mps: float = 3.0, length_km: float = 200.0) -> None:
    """Global wind: two zonal jets (peak m/s, latitude) of Gaussian
    meridional shape times sin(pi z / 24 km), plus Gaussian-correlated
    perturbations on u and v, all scaled by cos(latitude) so that the wind
    vanishes toward the poles; w = 0.  The noise is correlated on a 45
    degree-latitude cell width."""
    nx, ny, nz = st.nx, st.ny, st.nz
    zc = st.z_cc[2:nz + 2]
    lat = st.lat_cc[2:ny + 2]
    coslat = np.cos(lat * DEG2RAD)
    dx45 = st.dy * math.cos(45.0 * DEG2RAD) * (st.dx_deg / st.dn_deg)
    du = correlated_noise(nx, ny, dx45, st.dy, length_km, sigma_mps, rng)
    dv = correlated_noise(nx, ny, dx45, st.dy, length_km, sigma_mps, rng)
    prof = np.zeros(ny)
    for peak, lat0 in jets:
        prof += peak * np.exp(-((lat - lat0) / jet_width_deg) ** 2)
    shape_z = np.sin(np.pi * zc / 24.0)
    for kk in range(nz):
        st.vx_pd[2:nx + 2, 2:ny + 2, kk + 2] = ((prof * shape_z[kk])[None, :] + du) * coslat[None, :] * MPS_2_KMPHR
        st.vy_pd[2:nx + 2, 2:ny + 2, kk + 2] = dv * coslat[None, :] * MPS_2_KMPHR
    st.vz_pd[...] = 0.0


def global_state(seed: int = 5, ns: int = 20, nx: int = 1440, ny: int = 720, nz: int = 80,
                 bins=None, pole_gap_deg: float = 0.25) -> TransportState:
    """The large global run: a periodic 0.25 deg lon-lat grid of 1440x720
    cells (rows of 179.5/720 deg from 89.75 S to 89.75 N), 80 levels of
    0.3 km, 20 bins with |vf| log-spaced 0.005-8 m/s, two jets plus
    correlated noise vanishing toward the poles, Kh = 2000 m2/s,
    Kz = 10 m2/s, Crank-Nicolson, and a broad cloud aloft at 20-24 km (the
    top of the grid), where a 6 h umbrella-style source puts it.

    The rows stop pole_gap_deg short of each pole.  On a grid from exactly
    90 S to 90 N the ghost row beyond each pole gets the negative of the
    first row's volume from the spherical-shell formulas
    (src/Calc_Mesh.f90:121-139), so the average volume at the polar
    interface that diffCN_y and diff_y divide by
    (src/Diffusion.F90:1052-1058) is exactly zero and the reference's y
    diffusion returns NaN there.

    A smaller grid of the same construction comes from nx, ny, nz; bins
    selects a subset of the bins (indices into the ns log-spaced fall
    speeds), e.g. for a CPU sample."""
    vf_all = fall_velocities(ns, 0.005, 8.0)
    vf = vf_all if bins is None else vf_all[list(bins)]
    nb = len(vf)
    st = TransportState(nx, ny, nz, nb, limiter=LIM_SUPERBEE, use_cn=True, imp_fac=0.5)
    de, dn = 360.0 / nx, (180.0 - 2.0 * pole_gap_deg) / ny
    lonlat_mesh(st, -180.0, -90.0 + pole_gap_deg, de, dn, np.full(nz, 24.0 / nz))
    st.is_periodic = True
    st.dx_deg, st.dn_deg = de, dn
    rng = np.random.default_rng(seed)
    global_jets_winds(st, rng)
    set_uniform_fall(st, vf)
    set_diffusivity(st, 2000.0, 10.0)
    st.use_diffusion = True
    st.dt = adjust_dt(st, use_diffusion=True)
    # the broad cloud: centred at 15 N, 40 x 15 degrees, at 22 km
    cloud_window(st, (nx * 0.55, ny * (105.0 / 180.0), nz * (22.0 / 24.0)),
                 (nx * (10.0 / 360.0), ny * (4.0 / 180.0), nz * (1.5 / 24.0)), 1.0e3)
    return st


def inversion_case(seed: int = 4, nx: int = 240, ny: int = 160, nz: int = 50, ns: int = 8,
                   n_heights: int = 40, n_periods: int = 24, hours: float = 12.0):
    """The adjoint source-inversion workload: lon-lat 0.25 deg 240x160x50
    (20-60 N), dz 0.5 km, 8 bins (|vf| log-spaced 0.01-5 m/s, equal
    fractions), the forecast's zonal jet with seeded perturbations,
    Kh = Kz = 500 m2/s, Crank-Nicolson, an empty domain, a vent at grid
    cell (nx/4, ny/2).  The unknown source is a rate table of n_heights
    levels x n_periods periods over `hours` (30 min periods over 12 h); dt
    is the largest whole fraction of a period within the CFL / diffusion
    limit.  Returns (state, vent (i, j), steps per period, bin fractions,
    dz (nz), the true source table (a lognormal height profile centred at
    9 km, sigma_ln 0.35, times a Gaussian pulse in time))."""
    st = TransportState(nx, ny, nz, ns, limiter=LIM_SUPERBEE, use_cn=True, imp_fac=0.5)
    lonlat_mesh(st, 180.0, 20.0, 0.25, 0.25, np.full(nz, 0.5))
    rng = np.random.default_rng(seed)
    zonal_jet_winds(st, rng)
    set_uniform_fall(st, fall_velocities(ns, 0.01, 5.0))
    set_diffusivity(st, 500.0, 500.0)
    st.use_diffusion = True
    dt_max = adjust_dt(st, use_diffusion=True)
    period = hours / n_periods
    spp = int(math.ceil(period / dt_max))
    st.dt = period / spp
    frac = np.full(ns, 1.0 / ns)
    dz = np.full(nz, 0.5)
    zh = (np.arange(n_heights) + 0.5) * 0.5
    prof = np.exp(-(np.log(zh / 9.0)) ** 2 / (2 * 0.35 ** 2))
    tp = (np.arange(n_periods) + 0.5) * period
    env = np.exp(-0.5 * ((tp - 3.0) / 1.5) ** 2)
    s_true = 1.0e6 * np.outer(env, prof)
    return st, (nx // 4, ny // 2), spp, frac, dz, s_true


def random_state(nx, ny, nz, ns, seed, latlon=False, periodic=False, zscaling_id=0,
                 limiter=LIM_SUPERBEE, use_cn=True, imp_fac=0.5, diffusivity=True,
                 cfl=0.8) -> TransportState:
    """A small state with non-uniform, divergent winds, variable layer
    thickness, spatially variable K and a rough positive cloud: the parity
    test case generator."""
    rng = np.random.default_rng(seed)
    st = TransportState(nx, ny, nz, ns, limiter=limiter, use_cn=use_cn, imp_fac=imp_fac)
    dz = 0.5 + rng.random(nz)
    if latlon:
        de = 360.0 / nx if periodic else 0.5
        lonlat_mesh(st, 0.0, 30.0 + 10.0 * rng.random(), de, 0.5, dz)
        st.is_periodic = periodic
    else:
        cartesian_mesh(st, 3.0 + rng.random(), 3.0 + rng.random(), dz)
        st.is_periodic = periodic
    st.zscaling_id = zscaling_id
    P = (nx + 4, ny + 4, nz + 4)
    st.vx_pd[...] = (20.0 + 30.0 * rng.standard_normal(P)) * (1.0 if rng.random() < 0.5 else -1.0)
    st.vy_pd[...] = 25.0 * rng.standard_normal(P)
    st.vz_pd[...] = 2.0 * rng.standard_normal(P)
    st.vh_pd[...] = 0.5 * rng.standard_normal(P)
    vf = fall_velocities(ns, 0.05, 3.0)
    for n in range(ns):
        st.vf_pd[..., n] = vf[n] * (1.0 + 0.2 * rng.random(P))
    if diffusivity:
        st.kx[...] = 1.0 + rng.random(st.kx.shape)
        st.ky[...] = 1.0 + rng.random(st.ky.shape)
        st.kz[...] = 0.05 + 0.05 * rng.random(st.kz.shape)
    if zscaling_id == 2:
        st.deldxond_cc[...] = 0.05 * rng.standard_normal(st.deldxond_cc.shape)
        st.deldyond_cc[...] = 0.05 * rng.standard_normal(st.deldyond_cc.shape)
    # rough positive cloud with exact zeros in places (exercise the EPS skip)
    c = rng.random((nx, ny, nz, ns)) ** 3 * 100.0
    c[c < 5.0] = 0.0
    st.concen_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, :, 0] = c
    st.use_diffusion = diffusivity
    st.dt = adjust_dt(st, use_diffusion=diffusivity, cfl=cfl)
    return st
