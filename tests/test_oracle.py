"""Pins for the C oracle (the checker the HIP path is compared against).

The reference Fortran cannot be built here (its modules include a generated
header and need LAPACK and MetReader), so each oracle function is pinned by
what the mathematics fixes, chosen so that a plausible mistake fails:

* advection (advect_x/y/z): identity at zero velocity; exact one-cell shift
  at CFL = 1 in both directions (upwind side, update sign, outflow/deposit
  index); mass conservation incl. outflow with divergent winds and
  non-uniform cells (divergence correction, telescoping); total-variation
  and max/min bounds; agreement with an independently written flux-form
  implementation; textbook limiter values.
* diffusion: Neumann cosine eigenmodes with the closed-form amplification
  factor for explicit and theta-scheme sweeps; mass conservation with
  variable cells and K; theta = 0 against the explicit routine (a separate
  code path of the reference); first-order consistency of theta = 0.5 with
  the explicit operator on non-uniform grids.
* dgtsv: bitwise against scipy's LAPACK dgtsv, with and without pivoting.
* Set_BC: the copy rules, checked cell by cell on tiny grids.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from helpers import interior_mass, outflow_mass, rel_linf, total_variation


# --------------------------------------------------------------------------
# dgtsv
# --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 100])
@pytest.mark.parametrize("pivot", [False, True])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_dgtsv_matches_lapack(oracle, n, pivot, nrhs):
    rng = np.random.default_rng(100 * n + 10 * nrhs + pivot)
    dl = rng.standard_normal(max(n - 1, 0))
    du = rng.standard_normal(max(n - 1, 0))
    d = rng.standard_normal(n) * (0.1 if pivot else 1.0) + (0.0 if pivot else 4.0)
    b = rng.standard_normal((n, nrhs))
    x, info, *_ = oracle.dgtsv(dl, d, du, b)
    if n == 1:  # scipy's wrapper rejects empty off-diagonals
        assert info == 0
        np.testing.assert_array_equal(x, b / d[0])
        return
    ref = lapack.dgtsv(dl, d, du, b)
    assert info == ref[4] == 0
    np.testing.assert_array_equal(x, ref[3])


def test_dgtsv_singular_reports_info(oracle):
    x, info, *_ = oracle.dgtsv([0.0], [0.0, 1.0], [0.0], [1.0, 1.0])
    assert info == 1


# --------------------------------------------------------------------------
# limiter functions: textbook values of phi(theta)
# --------------------------------------------------------------------------
THETAS = [-1.0, 0.0, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0]
EXPECTED_PHI = {
    "superbee": [0, 0, 0.5, 1, 1, 1, 1.5, 2, 2],
    "minmod": [0, 0, 0.25, 0.5, 0.75, 1, 1, 1, 1],
    "mc": [0, 0, 0.5, 0.75, 0.875, 1, 1.25, 1.5, 2],
    "fromm": [0, 0.5, 0.625, 0.75, 0.875, 1, 1.25, 1.5, 2],
    "laxwen": [1] * 9,
    "bw": THETAS,
    "none": [0] * 9,
}


@pytest.mark.parametrize("name", list(EXPECTED_PHI))
def test_limiter_phi_textbook(oracle, pkg, name):
    lim = pkg.state.LIMITER_NAMES[name]
    got = [oracle.limiter_phi(lim, t) for t in THETAS]
    np.testing.assert_allclose(got, EXPECTED_PHI[name], rtol=0, atol=1e-15)


# --------------------------------------------------------------------------
# advection closed forms
# --------------------------------------------------------------------------
def _uniform_cart(pkg, nx, ny, nz, ns=1, dx=2.0, dy=2.0):
    st = pkg.TransportState(nx, ny, nz, ns)
    pkg.workloads.cartesian_mesh(st, dx, dy, np.ones(nz))
    rng = np.random.default_rng(7)
    st.concen_pd[2:nx + 2, 2:ny + 2, 2:nz + 2, :, 0] = rng.integers(0, 1000, (nx, ny, nz, ns))
    return st


@pytest.mark.parametrize("sweep", ["x", "y", "z"])
def test_advection_zero_velocity_is_identity(oracle, pkg, sweep):
    st = _uniform_cart(pkg, 9, 8, 7, 2)
    st.dt = 0.3
    c0 = st.interior().copy()
    getattr(oracle.Oracle(st), "advect_" + sweep)()
    np.testing.assert_array_equal(st.interior(), c0)
    np.testing.assert_array_equal(st.interior(1), c0)


@pytest.mark.parametrize("sweep", ["x", "y", "z"])
@pytest.mark.parametrize("sign", [1, -1])
def test_advection_cfl1_exact_shift(oracle, pkg, sweep, sign):
    """dt*|u|*sigma/kappa = 1: the limited DCU update is an exact one-cell
    shift and the cell leaving the domain goes into the outflow/deposit."""
    nx, ny, nz = 9, 8, 7
    st = _uniform_cart(pkg, nx, ny, nz)
    st.dt = 0.5  # kappa = 4, sigma = 2 (x,y) or 4 (z)
    if sweep == "x":
        st.vx_pd[...] = sign * 4.0
    elif sweep == "y":
        st.vy_pd[...] = sign * 4.0
    else:
        st.vf_pd[...] = sign * 2.0
    c0 = st.interior().copy()
    getattr(oracle.Oracle(st), "advect_" + sweep)()
    ax = {"x": 0, "y": 1, "z": 2}[sweep]
    expect = np.roll(c0, sign, axis=ax)
    sl = [slice(None)] * 4
    sl[ax] = 0 if sign > 0 else -1
    expect[tuple(sl)] = 0.0
    np.testing.assert_array_equal(st.interior(), expect)
    leaving = c0.take(-1 if sign > 0 else 0, axis=ax)
    if sweep == "x":
        of = (st.outflow_yz2_pd if sign > 0 else st.outflow_yz1_pd)[2:ny + 2, 2:nz + 2, :]
    elif sweep == "y":
        of = (st.outflow_xz2_pd if sign > 0 else st.outflow_xz1_pd)[2:nx + 2, 2:nz + 2, :]
    else:
        of = (st.outflow_xy2_pd if sign > 0 else st.outflow_xy1_pd)[2:nx + 2, 2:ny + 2, :]
    np.testing.assert_array_equal(of, leaving)
    if sweep == "z":
        np.testing.assert_array_equal(st.deposit_granularity, st.outflow_xy1_pd[2:nx + 2, 2:ny + 2, :])


def test_periodic_x_shift_wraps(oracle, pkg):
    nx, ny, nz = 9, 3, 2
    st = _uniform_cart(pkg, nx, ny, nz)
    st.is_periodic = True
    st.dt = 0.5
    st.vx_pd[...] = 4.0
    c0 = st.interior().copy()
    oracle.Oracle(st).advect_x()
    np.testing.assert_array_equal(st.interior(), np.roll(c0, 1, axis=0))
    assert not st.outflow_yz1_pd.any() and not st.outflow_yz2_pd.any()


@pytest.mark.parametrize("latlon", [False, True])
@pytest.mark.parametrize("sweep", ["x", "y", "z"])
def test_advection_conserves_mass_with_outflow(oracle, pkg, sweep, latlon):
    st = pkg.workloads.random_state(11, 10, 9, 3, seed=3 + latlon, latlon=latlon)
    m0 = interior_mass(st)
    getattr(oracle.Oracle(st), "advect_" + sweep)()
    m1 = interior_mass(st) + outflow_mass(st)
    np.testing.assert_allclose(m1, m0, rtol=1e-13)


def test_periodic_advection_conserves_mass(oracle, pkg):
    st = pkg.workloads.random_state(16, 6, 5, 2, seed=5, latlon=True, periodic=True)
    assert st.is_periodic
    m0 = interior_mass(st)
    oracle.Oracle(st).advect_x()
    np.testing.assert_allclose(interior_mass(st), m0, rtol=1e-13)


@pytest.mark.parametrize("limiter", ["superbee", "minmod", "mc"])
@pytest.mark.parametrize("cfl", [0.3, 0.8, 1.0])
def test_tvd_and_bounds_uniform_flow(oracle, pkg, limiter, cfl):
    nx = 64
    st = pkg.TransportState(nx, 2, 2, 1, limiter=pkg.state.LIMITER_NAMES[limiter])
    pkg.workloads.cartesian_mesh(st, 2.0, 2.0, np.ones(2))
    rng = np.random.default_rng(11)
    row = np.where(rng.random(nx) < 0.3, 0.0, rng.random(nx) * 10)
    st.concen_pd[2:nx + 2, 2:4, 2:4, 0, 0] = row[:, None, None]
    st.vx_pd[...] = 4.0
    st.dt = cfl * 0.5
    oracle.Oracle(st).advect_x()
    new = st.interior()[:, 0, 0, 0]
    tv0 = total_variation(np.concatenate([[0.0], row, [0.0]]))
    tv1 = total_variation(np.concatenate([[0.0], new, [0.0]]))
    assert tv1 <= tv0 * (1 + 1e-14)
    assert new.min() >= -1e-12 and new.max() <= row.max() * (1 + 1e-14)


def _flux_form_row(q, vel, sig, kap, dt, limiter):
    """Independent vectorised statement of one full-row sweep in conservative
    flux form q_i -= dt/kappa_i (F_{i+1/2} - F_{i-1/2}), F = upwind flux plus
    the limited correction 1/2|a|(1 - dt|a|/kappa)phi(theta)dq.  Arrays hold
    cells -1..N+2; returns the update on cells 0..N+1."""
    N = q.size - 4
    a = np.zeros(N + 5)  # interface I at index I+1, I = -1..N+3
    I = np.arange(0, N + 2)
    a[I + 1] = 0.5 * (vel[I - 1 + 1] + vel[I + 1]) * sig[I + 1]
    F = np.zeros(N + 5)
    J = np.arange(0, N + 3)
    F[J + 1] = np.maximum(a[J + 1], 0) * q[J - 1 + 1] + np.minimum(a[J + 1], 0) * q[J + 1]
    dq = np.zeros(N + 5)
    K = np.arange(0, N + 3)
    dq[K + 1] = q[K + 1] - q[K - 1 + 1]
    fss = np.zeros(N + 5)
    for Ii in range(1, N + 2):
        d = dq[Ii + 1]
        if abs(d) <= 1e-10:
            continue
        du = dq[Ii] if a[Ii + 1] > 0 else dq[Ii + 2]
        th = np.sign(du * d) * 3.0 if abs(du) > 3 * abs(d) else du / d
        phi = {"superbee": max(0.0, min(1.0, 2 * th), min(2.0, th)),
               "minmod": max(0.0, min(1.0, th)),
               "mc": max(0.0, min((1 + th) / 2, 2.0, 2 * th)),
               "fromm": 0.5 * (1 + th), "laxwen": 1.0}[limiter]
        aa = abs(a[Ii + 1])
        fss[Ii + 1] = 0.5 * aa * (1 - dt / kap[Ii + 1] * aa) * d * phi
    c = np.arange(0, N + 2)
    return -dt / kap[c + 1] * (F[c + 2] - F[c + 1] + fss[c + 2] - fss[c + 1])


@pytest.mark.parametrize("limiter", ["superbee", "minmod", "mc", "fromm", "laxwen"])
def test_row_matches_flux_form(oracle, pkg, limiter):
    st = pkg.workloads.random_state(23, 4, 3, 1, seed=21, latlon=True,
                                    limiter=pkg.state.LIMITER_NAMES[limiter])
    before = st.copy()
    oracle.Oracle(before).Set_BC(1)
    oracle.Oracle(st).advect_y()
    nx, ny = st.nx, st.ny
    for i in range(2, nx + 2):
        for k in range(2, st.nz + 2):
            q = before.concen_pd[i, :, k, 0, 0]
            upd = _flux_form_row(q, before.vy_pd[i, :, k], before.sigma_ny_pd[i, :, k],
                                 before.kappa_pd[i, :, k], st.dt, limiter)
            got = st.concen_pd[i, 2:ny + 2, k, 0, 0] - q[2:ny + 2]
            assert np.max(np.abs(got - upd[1:ny + 1])) <= 1e-12 * np.max(np.abs(q))
            np.testing.assert_allclose(st.outflow_xz1_pd[i, k, 0], upd[0], rtol=1e-12, atol=1e-12 * q.max())
            np.testing.assert_allclose(st.outflow_xz2_pd[i, k, 0], upd[ny + 1], rtol=1e-12,
                                       atol=1e-12 * q.max())


def test_mirror_symmetry_x(oracle, pkg):
    """Reversing x and the wind gives the mirrored result (uniform cells)."""
    st = _uniform_cart(pkg, 17, 3, 2)
    rng = np.random.default_rng(4)
    st.vx_pd[...] = 3.0 + rng.random(st.vx_pd.shape)
    st.dt = 0.2
    mir = st.copy()
    mir.concen_pd[...] = st.concen_pd[::-1]
    mir.vx_pd[...] = -st.vx_pd[::-1]
    oracle.Oracle(st).advect_x()
    oracle.Oracle(mir).advect_x()
    assert rel_linf(mir.interior()[::-1], st.interior()) < 1e-14
    np.testing.assert_allclose(mir.outflow_yz1_pd, st.outflow_yz2_pd, rtol=1e-13)


def test_advect_horz_order_by_parity(oracle, pkg):
    base = pkg.workloads.random_state(10, 9, 4, 2, seed=8)
    a, b, c, d = base.copy(), base.copy(), base.copy(), base.copy()
    oracle.Oracle(a).AdvectHorz(2)
    ob = oracle.Oracle(b)
    ob.advect_y()
    ob.advect_x()
    np.testing.assert_array_equal(a.concen_pd, b.concen_pd)
    oracle.Oracle(c).AdvectHorz(3)
    od = oracle.Oracle(d)
    od.advect_x()
    od.advect_y()
    np.testing.assert_array_equal(c.concen_pd, d.concen_pd)
    assert not np.array_equal(a.interior(), c.interior())


def test_not_aloft_bins_are_zeroed(oracle, pkg):
    st = pkg.workloads.random_state(8, 7, 6, 3, seed=9)
    st.is_aloft = np.array([1, 0, 1], dtype=np.int32)
    oracle.Oracle(st).advect_x()
    assert not st.interior()[..., 1].any()
    assert st.interior()[..., 0].any() and st.interior()[..., 2].any()


# --------------------------------------------------------------------------
# Set_BC
# --------------------------------------------------------------------------
@pytest.mark.parametrize("periodic", [False, True])
def test_set_bc1_rules(oracle, pkg, periodic):
    st = pkg.workloads.random_state(6, 5, 4, 2, seed=12)
    st.is_periodic = periodic
    rng = np.random.default_rng(0)
    st.concen_pd[...] = rng.random(st.concen_pd.shape)
    ref = st.copy()
    oracle.Oracle(st).Set_BC(1)
    nx, ny, nz = st.nx, st.ny, st.nz
    c, c0 = st.concen_pd[..., 0], ref.concen_pd[..., 0]
    # concentrations: interior untouched, ghosts zero except periodic x copies
    np.testing.assert_array_equal(c[2:nx + 2, 2:ny + 2, 2:nz + 2], c0[2:nx + 2, 2:ny + 2, 2:nz + 2])
    ghost = np.ones(c.shape[:3], bool)
    ghost[2:nx + 2, 2:ny + 2, 2:nz + 2] = False
    expect = np.zeros_like(c)
    if periodic:
        for gi, si in ((0, nx - 1), (1, nx), (nx + 2, 1), (nx + 3, 2)):
            expect[gi, 2:ny + 2, 2:nz + 2] = c0[si + 1, 2:ny + 2, 2:nz + 2]
    np.testing.assert_array_equal(c[ghost], expect[ghost])
    np.testing.assert_array_equal(st.concen_pd[..., 1], ref.concen_pd[..., 1])
    # velocities: zero-order extrapolation, periodic wrap in x
    vx, vx0 = st.vx_pd, ref.vx_pd
    src = (nx - 1, nx, 1, 2) if periodic else (1, 1, nx, nx)
    for gi, si in zip((0, 1, nx + 2, nx + 3), src):
        np.testing.assert_array_equal(vx[gi], vx0[si + 1])
    for gj, sj in zip((0, 1, ny + 2, ny + 3), (1, 1, ny, ny)):
        np.testing.assert_array_equal(st.vy_pd[:, gj], ref.vy_pd[:, sj + 1])
    for gk, sk in zip((0, 1, nz + 2, nz + 3), (1, 1, nz, nz)):
        np.testing.assert_array_equal(st.vz_pd[:, :, gk], ref.vz_pd[:, :, sk + 1])
        np.testing.assert_array_equal(st.vf_pd[:, :, gk], ref.vf_pd[:, :, sk + 1])
    np.testing.assert_array_equal(st.vh_pd, ref.vh_pd)


@pytest.mark.parametrize("periodic", [False, True])
def test_set_bc2_neumann_mirror(oracle, pkg, periodic):
    """Sequential x (periodic override), y, z mirrors leave every ghost cell
    equal to the interior cell its clamped (or x-wrapped) index names."""
    st = pkg.workloads.random_state(6, 5, 4, 2, seed=13)
    st.is_periodic = periodic
    rng = np.random.default_rng(1)
    st.concen_pd[...] = rng.random(st.concen_pd.shape)
    ref = st.copy()
    oracle.Oracle(st).Set_BC(2)
    nx, ny, nz = st.nx, st.ny, st.nz

    def mx(i):
        if 1 <= i <= nx:
            return i
        if periodic:
            return {-1: nx - 1, 0: nx, nx + 1: 1, nx + 2: 2}[i]
        return 1 if i < 1 else nx

    for i in range(-1, nx + 3):
        for j in range(-1, ny + 3):
            for k in range(-1, nz + 3):
                jj, kk = min(max(j, 1), ny), min(max(k, 1), nz)
                exp = ref.concen_pd[mx(i) + 1, jj + 1, kk + 1, :, 0]
                np.testing.assert_array_equal(st.concen_pd[i + 1, j + 1, k + 1, :, 0], exp)


def test_set_bc_bad_code(oracle, pkg):
    st = pkg.workloads.random_state(4, 4, 4, 1, seed=1)
    assert oracle.Oracle(st).Set_BC(3) != 0


# --------------------------------------------------------------------------
# diffusion
# --------------------------------------------------------------------------
def _mode_state(pkg, dirn, N, m, K, dt, use_cn, imp_fac=0.5):
    shape = [3, 3, 3]
    shape[dirn] = N
    st = pkg.TransportState(*shape, 1, use_cn=use_cn, imp_fac=imp_fac)
    d = [2.0, 3.0, 0.5]
    pkg.workloads.cartesian_mesh(st, d[0], d[1], np.full(shape[2], d[2]))
    st.kx[...] = K
    st.ky[...] = K
    st.kz[...] = K
    st.dt = dt
    l = np.arange(1, N + 1)
    mode = np.cos(np.pi * m * (l - 0.5) / N)
    sl = [None, None, None]
    sl[dirn] = slice(None)
    st.concen_pd[2:shape[0] + 2, 2:shape[1] + 2, 2:shape[2] + 2, 0, 0] = 5.0 + 0 * st.interior()[..., 0]
    st.concen_pd[2:shape[0] + 2, 2:shape[1] + 2, 2:shape[2] + 2, 0, 0] += mode[tuple(sl)]
    r = K * dt / d[dirn] ** 2
    return st, mode, r


@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("method", ["explicit", "cn", "be"])
@pytest.mark.parametrize("m", [1, 3])
def test_diffusion_eigenmode_decay(oracle, pkg, dirn, method, m):
    N = 20
    imp = {"explicit": 0.0, "cn": 0.5, "be": 1.0}[method]
    st, mode, r = _mode_state(pkg, dirn, N, m, K=0.7, dt=0.3, use_cn=(method != "explicit"), imp_fac=imp)
    s2 = np.sin(np.pi * m / (2 * N)) ** 2
    g = (1 - (1 - imp) * 4 * r * s2) / (1 + imp * 4 * r * s2)
    O = oracle.Oracle(st)
    name = ("diff_" if method == "explicit" else "diffCN_") + "xyz"[dirn]
    getattr(O, name)()
    got = np.moveaxis(st.interior()[..., 0], dirn, 0)
    exp = 5.0 + g * mode
    assert np.max(np.abs(got - exp[:, None, None])) < 1e-13 * 6


@pytest.mark.parametrize("dirn", [0, 1, 2])
@pytest.mark.parametrize("use_cn", [False, True])
def test_diffusion_conserves_mass(oracle, pkg, dirn, use_cn):
    st = pkg.workloads.random_state(12, 11, 10, 2, seed=30 + dirn, latlon=True, use_cn=use_cn)
    st.dt = st.dt * (0.5 if not use_cn else 20.0)
    m0 = interior_mass(st)
    O = oracle.Oracle(st)
    O.Set_BC(2)
    getattr(O, ("diffCN_" if use_cn else "diff_") + "xyz"[dirn])()
    np.testing.assert_allclose(interior_mass(st), m0, rtol=1e-13)


@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_cn_theta0_equals_explicit(oracle, pkg, dirn):
    st = pkg.workloads.random_state(10, 9, 8, 2, seed=40 + dirn, latlon=True, imp_fac=0.0)
    ex = st.copy()
    ex.use_cn = False
    getattr(oracle.Oracle(st), "diffCN_" + "xyz"[dirn])()
    getattr(oracle.Oracle(ex), "diff_" + "xyz"[dirn])()
    assert rel_linf(st.interior(), ex.interior()) < 1e-14


@pytest.mark.parametrize("dirn", [0, 1, 2])
def test_cn_consistent_with_explicit_operator(oracle, pkg, dirn):
    """On a non-uniform grid, CN(dt) - explicit(dt) = O(dt^2): halving dt
    divides the gap by ~4."""
    gaps = []
    for scale in (1.0, 0.5):
        st = pkg.workloads.random_state(10, 9, 8, 1, seed=50 + dirn, latlon=True)
        st.dt = 0.02 * scale
        ex = st.copy()
        ex.use_cn = False
        getattr(oracle.Oracle(st), "diffCN_" + "xyz"[dirn])()
        getattr(oracle.Oracle(ex), "diff_" + "xyz"[dirn])()
        gaps.append(np.max(np.abs(st.interior() - ex.interior())))
    assert 3.5 < gaps[0] / gaps[1] < 4.5


def test_cn_single_cell_line_copies_nothing(oracle, pkg):
    st = pkg.workloads.random_state(5, 4, 3, 1, seed=60)
    st.bounds = (2, 2, 1, 4, 1, 3)  # one cell along x
    c0 = st.interior().copy()
    oracle.Oracle(st).diffCN_x()
    np.testing.assert_array_equal(st.interior(), c0)
    assert not st.concen_pd[..., 1].any()


# --------------------------------------------------------------------------
# the full step
# --------------------------------------------------------------------------
@pytest.mark.parametrize("use_diffusion", [False, True])
@pytest.mark.parametrize("periodic", [False, True])
def test_step_conserves_total_mass(oracle, pkg, use_diffusion, periodic):
    st = pkg.workloads.random_state(16, 9, 8, 3, seed=70, latlon=True, periodic=periodic,
                                    diffusivity=use_diffusion)
    O = oracle.Oracle(st)
    m0 = interior_mass(st)
    for itime in (1, 2, 3):
        O.transport_step(itime, use_diffusion)
        O.step_glue()
    # the glue moved the result to ts0 and cleared ts1
    np.testing.assert_allclose(interior_mass(st) + outflow_mass(st), m0, rtol=1e-13)


def test_lonlat_shell_volume(pkg):
    """Global lon-lat cells tile the spherical shell: sum(kappa) over a layer
    is 4/3 pi (r2^3 - r1^3)."""
    st = pkg.TransportState(36, 18, 3, 1)
    pkg.workloads.lonlat_mesh(st, 0.0, -90.0, 10.0, 10.0, np.array([1.0, 2.0, 0.5]))
    assert st.is_periodic
    z = np.concatenate([[0.0], np.cumsum([1.0, 2.0, 0.5])])
    R = pkg.workloads.RAD_EARTH
    for k in range(3):
        vol = st.kappa_pd[2:38, 2:20, k + 2].sum()
        exp = 4.0 / 3.0 * np.pi * ((R + z[k + 1]) ** 3 - (R + z[k]) ** 3)
        np.testing.assert_allclose(vol, exp, rtol=1e-12)


def test_pole_to_pole_grid_y_diffusion_is_nan(pkg, oracle):
    """On a lon-lat grid from exactly 90 S to 90 N the ghost row beyond
    each pole gets the negated volume of the first row
    (src/Calc_Mesh.f90:121-139), so the average volume diffCN_y divides by
    at the polar interface (src/Diffusion.F90:1052-1058) is zero and the
    reference's y diffusion returns NaN in the polar rows; the global
    workload's rows stop 0.25 deg short of the poles and stay finite."""
    st = pkg.workloads.global_state(nx=8, ny=720, nz=2, ns=1, pole_gap_deg=0.0)
    kap = st.kappa_pd[4, :, 3]
    assert kap[1] == -kap[2] and kap[-2] == -kap[-3]  # ghost row 0 = -row 1 (and at the north pole)
    oracle.Oracle(st).diffCN_y()
    q = st.concen_pd[2:10, 2:722, 2:4, 0, 0]
    assert np.isnan(q[:, 0]).all() and np.isnan(q[:, -1]).all()
    st = pkg.workloads.global_state(nx=8, ny=720, nz=2, ns=1)
    oracle.Oracle(st).diffCN_y()
    assert np.isfinite(st.concen_pd).all()
