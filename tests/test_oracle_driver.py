"""Oracle pins for the driver work around the hot path: wind time
interpolation (src/MesoInterpolater.f90:203-230), Adjust_DT
(src/Adjust_dt.f90:89-333) and get_minmax_index / the FAST_SUBGRID step
(src/Calc_Mesh.f90:432-542, src/AdvectionHorz.f90:60-66).

Each expectation below is derived by hand from the reference's definition
(exact endpoints of the interpolation, closed-form CFL limits on uniform
grids, bounding boxes of hand-placed clouds), not by re-running the
oracle's formula."""
import numpy as np
import pytest


def f32(shape, seed, lo=-30.0, hi=30.0):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(lo, hi, shape).astype(np.float32))


def interior(st, a):
    return a[2:st.nx + 2, 2:st.ny + 2, 2:st.nz + 2]


def met_pair(st, seed):
    sh, shf = (st.nx, st.ny, st.nz), (st.nx, st.ny, st.nz, st.ns)
    last = (f32(sh, seed), f32(sh, seed + 1), f32(sh, seed + 2, -1, 1), f32(shf, seed + 3, -5, 0))
    nxt = (f32(sh, seed + 4), f32(sh, seed + 5), f32(sh, seed + 6, -1, 1), f32(shf, seed + 7, -5, 0))
    return last, nxt


def test_met_interp_endpoints_and_midpoint(pkg, oracle):
    st = pkg.workloads.random_state(7, 6, 5, 2, seed=1)
    # integer m/s: every FP32 difference and every product below is exact
    sh, shf = (7, 6, 5), (7, 6, 5, 2)
    rng = np.random.default_rng(2)
    ints = lambda s: np.asfortranarray(rng.integers(-40, 40, s).astype(np.float32))
    last = (ints(sh), ints(sh), ints(sh), ints(shf))
    nxt = (ints(sh), ints(sh), ints(sh), ints(shf))
    before = st.vx_pd.copy()
    O = oracle.Oracle(st)
    O.met_interp(last, nxt, 0.0, True)
    assert np.array_equal(interior(st, st.vx_pd), last[0].astype(np.float64) * 3.6)
    assert np.array_equal(interior(st, st.vf_pd), last[3].astype(np.float64) * 3.6)
    O.met_interp(last, nxt, 1.0, True)
    assert np.array_equal(interior(st, st.vy_pd), nxt[1].astype(np.float64) * 3.6)
    O.met_interp(last, nxt, 0.5, False)
    mid = (last[2].astype(np.float64) + nxt[2].astype(np.float64)) / 2.0
    assert np.allclose(interior(st, st.vz_pd), mid * 3.6, rtol=1e-15, atol=0)
    # ghosts are not touched
    g = st.vx_pd.copy()
    interior(st, g)[...] = 0.0
    gb = before.copy()
    interior(st, gb)[...] = 0.0
    assert np.array_equal(g, gb)


def test_met_interp_difference_in_single_precision(pkg, oracle):
    st = pkg.workloads.random_state(3, 3, 3, 1, seed=1)
    sh = (3, 3, 3)
    last = [np.asfortranarray(np.full(sh, 1.0, np.float32)) for _ in range(3)] + [None]
    nxt = [np.asfortranarray(np.full(sh, 1.0 + 2.0 ** -20, np.float32)) for _ in range(3)] + [None]
    nxt[0][...] = np.float32(16777217.0)  # 2^24 + 1 rounds to 2^24 in FP32
    last[0][...] = np.float32(1.0)
    oracle.Oracle(st).met_interp(last, nxt, 1.0, False)
    d32 = np.float32(16777216.0) - np.float32(1.0)  # FP32 difference: 16777215
    assert np.all(interior(st, st.vx_pd) == (1.0 + float(d32)) * 3.6)


def dt_params(oracle, **kw):
    p = oracle.DtParams()
    base = dict(is_latlon=0, use_met_next=0, n_gs_aloft=1, use_diffusion=0, use_cn=1, dx=5.0, dy=5.0,
                imp_dt_fac=4.0, cfl=0.8, dt_min=1e-5, dt_max=1.0, eps_small=1e-6, time=0.0,
                next_write_time=100.0, simtime=100.0)
    base.update(kw)
    for k, v in base.items():
        if k != "dz":
            setattr(p, k, v)
    return p


def uniform_cartesian(pkg, nx=8, ny=7, nz=6, ns=2, u=10.0, v=5.0, vf=-0.5):
    st = pkg.workloads.random_state(nx, ny, nz, ns, seed=3)
    pkg.workloads.cartesian_mesh(st, 5.0, 5.0, np.ones(nz))
    st.vx_pd[...] = u * 3.6
    st.vy_pd[...] = v * 3.6
    st.vz_pd[...] = 0.0
    st.vf_pd[...] = vf * 3.6
    return st


def test_adjust_dt_cfl_closed_form(pkg, oracle):
    st = uniform_cartesian(pkg)
    dz = np.ones(st.nz)
    p = dt_params(oracle)
    p.dz = dz.ctypes.data_as(oracle._dp)
    dt, diag = oracle.Oracle(st).adjust_dt(p)
    # |u|/dx = 36/5 = 7.2 1/h dominates |v|/dy = 3.6 and |vf|/dz = 1.8
    assert diag[0] == 36.0 / 5.0 and diag[1] == 18.0 / 5.0 and diag[2] == 1.8
    expect = 1e-5 * int((0.8 * (1.0 / 7.2)) / 1e-5)
    assert dt == expect
    # no ash aloft: the vertical term is dropped
    p.n_gs_aloft = 0
    assert oracle.Oracle(st).adjust_dt(p)[1][2] == 0.0


def test_adjust_dt_output_and_stop_time_clips(pkg, oracle):
    st = uniform_cartesian(pkg)
    dz = np.ones(st.nz)
    p = dt_params(oracle, time=2.0, next_write_time=2.05)
    p.dz = dz.ctypes.data_as(oracle._dp)
    assert oracle.Oracle(st).adjust_dt(p)[0] == 2.05 - 2.0
    p = dt_params(oracle, time=2.0, simtime=2.07)
    p.dz = dz.ctypes.data_as(oracle._dp)
    assert oracle.Oracle(st).adjust_dt(p)[0] == 2.07 - 2.0


def test_adjust_dt_diffusion_limit(pkg, oracle):
    st = uniform_cartesian(pkg, u=0.1, v=0.1, vf=-0.01)
    st.kx[...], st.ky[...], st.kz[...] = 100.0, 100.0, 10.0
    dz = np.ones(st.nz)
    for use_cn, fac in ((1, 4.0), (0, 0.5)):
        p = dt_params(oracle, use_diffusion=1, use_cn=use_cn)
        p.dz = dz.ctypes.data_as(oracle._dp)
        _, diag = oracle.Oracle(st).adjust_dt(p)
        # Cartesian cells: kappa/sigma_x = dx, kappa/sigma_z = dz; below DT_MAX = 1 h
        assert diag[3] == (25.0 / 100.0 + 25.0 / 100.0 + 1.0 / 10.0) * fac


def test_adjust_dt_latlon_cellwise(pkg, oracle):
    st = pkg.workloads.random_state(9, 8, 5, 3, seed=11, latlon=True)
    dz = 0.5 + np.arange(5) * 0.1
    p = dt_params(oracle, is_latlon=1)
    p.dz = dz.ctypes.data_as(oracle._dp)
    _, diag = oracle.Oracle(st).adjust_dt(p)
    # brute force over cells with the reference's expressions
    best = [0.0, 0.0, 0.0]
    P = lambda a, i, j, k: a[i + 1, j + 1, k + 1]
    for i in range(1, 10):
        for j in range(1, 9):
            for k in range(1, 6):
                ms = min(P(st.sigma_nx_pd, i, j, k), P(st.sigma_nx_pd, i + 1, j, k))
                best[0] = max(best[0], abs(P(st.vx_pd, i, j, k)) * ms / P(st.kappa_pd, i, j, k))
                ms = min(P(st.sigma_ny_pd, i, j, k), P(st.sigma_ny_pd, i, j + 1, k))
                best[1] = max(best[1], abs(P(st.vy_pd, i, j, k)) * ms / P(st.kappa_pd, i, j, k))
                vfm = max(abs(st.vf_pd[i + 1, j + 1, k + 1, n]) for n in range(3))
                best[2] = max(best[2], (abs(P(st.vz_pd, i, j, k)) + vfm) / dz[k - 1])
    assert list(diag[:3]) == best


def place(st, box, value=1.0):
    i0, i1, j0, j1, k0, k1 = box
    st.concen_pd[...] = 0.0
    st.concen_pd[i0 + 1:i1 + 2, j0 + 1:j1 + 2, k0 + 1:k1 + 2, 0, 0] = value


def test_minmax_index_boxes(pkg, oracle):
    st = pkg.workloads.random_state(20, 16, 12, 2, seed=5)
    # cloud in cells i 6..9, j 5..8, k 3..6; vent (8,6,4)
    place(st, (6, 9, 5, 8, 3, 6))
    rc, b = oracle.Oracle(st).minmax_index(8, 6, 4)
    # the scan from below (up to the vent) stops at the first slab with ash
    # and keeps the last empty one; from the top likewise
    assert rc == 0 and b == (5, 10, 4, 9, 2, 7)
    # below the threshold counts as empty
    place(st, (6, 9, 5, 8, 3, 6), value=0.5e-7)
    rc, b = oracle.Oracle(st).minmax_index(8, 6, 4)
    assert b == (7, 9, 5, 7, 3, 5)  # every slab empty: boxes of 3 cells around the vents
    # empty domain, vent at (10, 8, 6)
    st.concen_pd[...] = 0.0
    rc, b = oracle.Oracle(st).minmax_index(10, 8, 6)
    assert b == (9, 11, 7, 9, 5, 7)
    # vent in the last cells: the scans leave 2-cell boxes, widened to 3
    rc, b = oracle.Oracle(st).minmax_index(20, 16, 12)
    assert b == (18, 20, 14, 16, 10, 12)


def test_minmax_index_periodic_touching(pkg, oracle):
    st = pkg.workloads.random_state(20, 16, 12, 1, seed=6, latlon=True, periodic=True)
    place(st, (1, 3, 5, 8, 3, 6))
    rc, b = oracle.Oracle(st).minmax_index(2, 6, 4)
    assert b[:2] == (1, 20)
    st.is_periodic = False
    rc, b = oracle.Oracle(st).minmax_index(2, 6, 4)
    assert b[:2] == (1, 4)


@pytest.mark.parametrize("itime", [1, 2])
def test_subgrid_step_equals_full_step_without_diffusion(pkg, oracle, itime):
    """With no diffusion and all ash inside the box the FAST_SUBGRID step
    changes nothing inside the box (the box keeps an empty margin, so no
    stencil reaches a cell it zeroes)."""
    st = pkg.workloads.random_state(24, 20, 12, 2, seed=9, latlon=True)
    c = st.concen_pd[..., 0].copy()
    st.concen_pd[...] = 0.0
    st.concen_pd[8:16, 7:14, 4:9, :, 0] = c[8:16, 7:14, 4:9, :]
    a, b = st.copy(), st.copy()
    oracle.Oracle(a).transport_step(itime, False)
    rc, box = oracle.Oracle(b).transport_step_subgrid(itime, False, 10, 9, 5)
    assert rc == 0 and box != (1, 24, 1, 20, 1, 12)
    assert np.array_equal(interior(a, a.concen_pd[..., 0]), interior(b, b.concen_pd[..., 0]))
