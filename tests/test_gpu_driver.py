"""Device driver work (wind time interpolation, Adjust_DT, get_minmax_index,
the FAST_SUBGRID step) through the C ABI against the oracle: bitwise."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARRAYS = ("concen_pd", "vx_pd", "vy_pd", "vz_pd", "vh_pd", "vf_pd", "outflow_xz1_pd", "outflow_xz2_pd",
          "outflow_yz1_pd", "outflow_yz2_pd", "outflow_xy1_pd", "outflow_xy2_pd", "deposit_granularity")


@pytest.fixture(scope="module")
def tr(pkg):
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def f32(shape, seed, lo, hi):
    return np.asfortranarray(np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32))


def met_steps(st, seed):
    sh, shf = (st.nx, st.ny, st.nz), (st.nx, st.ny, st.nz, st.ns)
    return [(f32(sh, seed + 10 * s, -30, 30), f32(sh, seed + 10 * s + 1, -30, 30),
             f32(sh, seed + 10 * s + 2, -1, 1), f32(shf, seed + 10 * s + 3, -5, 0)) for s in range(3)]


def same(a, b, names=ARRAYS):
    for n in names:
        x, y = getattr(a, n), getattr(b, n)
        assert np.array_equal(x, y), f"{n}: max diff {np.abs(x - y).max()}"


@pytest.mark.parametrize("use_vf", [True, False])
def test_met_interp(pkg, oracle, tr, use_vf):
    st = pkg.workloads.random_state(13, 10, 7, 3, seed=2, latlon=True)
    steps = met_steps(st, 5)
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    T.met_push(*steps[0])
    T.met_push(*steps[1])
    for frac in (0.0, 0.3, 0.999):
        O.met_interp(steps[0], steps[1], frac, use_vf)
        T.met_interp(frac, use_vf)
        T.sync_to_host()
        same(a, b)
    T.met_push(*steps[2])  # next -> last
    O.met_interp(steps[1], steps[2], 0.25, use_vf)
    T.met_interp(0.25, use_vf)
    T.sync_to_host()
    T.close()
    same(a, b)


def test_steps_follow_wind_changes(pkg, oracle, tr):
    """The library applies Set_BC(1) to the velocities only when they have
    changed since its last pass; steps between met interpolations and a
    host upload of new winds and fall velocities match the oracle, which
    applies it every step, bitwise."""
    st = pkg.workloads.random_state(14, 10, 7, 3, seed=8, latlon=True)
    st.concen_pd[..., 1] = st.concen_pd[..., 0]
    steps = met_steps(st, 9)
    a, b = st.copy(), st.copy()
    O, T = oracle.Oracle(a), tr.Transport(b)
    T.met_push(*steps[0])
    T.met_push(*steps[1])
    itime = 0
    for frac in (0.0, 0.5):
        O.met_interp(steps[0], steps[1], frac, True)
        T.met_interp(frac, True)
        for _ in range(2):
            itime += 1
            for X in (O, T):
                X.step_glue()
                X.transport_step(itime, True)
    rng = np.random.default_rng(4)
    new = {n: 20.0 * rng.standard_normal(getattr(a, n).shape) for n in ("vx_pd", "vy_pd", "vz_pd")}
    new["vf_pd"] = -np.abs(rng.standard_normal(a.vf_pd.shape))
    for h in (a, b):
        for n, v in new.items():
            getattr(h, n)[...] = v
    T.sync_from_host(tr.SYNC_WINDS | tr.SYNC_VF)
    for _ in range(2):
        itime += 1
        for X in (O, T):
            X.step_glue()
            X.transport_step(itime, True)
    T.sync_to_host()
    T.close()
    same(a, b)


def params(cls, dz, **kw):
    p = cls()
    base = dict(is_latlon=0, use_met_next=0, n_gs_aloft=3, use_diffusion=0, use_cn=1, dx=4.0, dy=3.0,
                imp_dt_fac=4.0, cfl=0.8, dt_min=1e-5, dt_max=1.0, eps_small=1e-6, time=1.0,
                next_write_time=5.0, simtime=50.0)
    base.update(kw)
    for k, v in base.items():
        setattr(p, k, v)
    p.dz = dz.ctypes.data_as(cls._fields_[7][1])
    return p


@pytest.mark.parametrize("latlon", [False, True])
@pytest.mark.parametrize("met", [0, 1])
@pytest.mark.parametrize("diff", [(0, 1), (1, 1), (1, 0)])
def test_adjust_dt(pkg, oracle, tr, latlon, met, diff):
    st = pkg.workloads.random_state(15, 12, 8, 3, seed=7, latlon=latlon)
    dz = 0.4 + 0.1 * np.arange(8)
    steps = met_steps(st, 9)
    kw = dict(is_latlon=int(latlon), use_met_next=met, use_diffusion=diff[0], use_cn=diff[1])
    dto, diago = oracle.Oracle(st.copy()).adjust_dt(params(oracle.DtParams, dz, **kw), steps[1])
    T = tr.Transport(st.copy())
    T.met_push(*steps[0])
    T.met_push(*steps[1])
    dtg, diagg = T.adjust_dt(params(tr.DtParams, dz, **kw))
    T.close()
    assert dtg == dto and np.array_equal(diagg, diago), (dtg, dto, diagg, diago)


def test_adjust_dt_clips_and_no_ash(pkg, oracle, tr):
    st = pkg.workloads.random_state(9, 8, 6, 2, seed=3)
    dz = np.ones(6)
    T = tr.Transport(st.copy())
    for kw in (dict(time=4.99), dict(simtime=1.001), dict(n_gs_aloft=0)):
        dto, diago = oracle.Oracle(st.copy()).adjust_dt(params(oracle.DtParams, dz, **kw))
        dtg, diagg = T.adjust_dt(params(tr.DtParams, dz, **kw))
        assert dtg == dto and np.array_equal(diagg, diago)
    T.close()


@pytest.mark.parametrize("seed", range(4))
def test_minmax_index(pkg, oracle, tr, seed):
    rng = np.random.default_rng(seed)
    st = pkg.workloads.random_state(30, 25, 14, 3, seed=seed, latlon=True, periodic=bool(seed % 2))
    c = st.concen_pd[..., 0].copy()
    st.concen_pd[...] = 0.0
    i0, j0, k0 = (int(v) for v in rng.integers(3, 10, 3))
    st.concen_pd[i0:i0 + 12, j0:j0 + 9, k0:k0 + 4, :, 0] = c[i0:i0 + 12, j0:j0 + 9, k0:k0 + 4, :] * 1e-6
    vent = (i0 + 3, j0 + 3, k0 + 1)
    rc, bo = oracle.Oracle(st.copy()).minmax_index(*vent)
    T = tr.Transport(st.copy())
    bg = T.minmax_index(*vent)
    T.close()
    assert rc == 0 and bg == bo


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("use_cn", [True, False])
def test_step_subgrid(pkg, oracle, tr, periodic, use_cn):
    st = pkg.workloads.random_state(28, 22, 12, 3, seed=4, latlon=True, periodic=periodic, use_cn=use_cn)
    if not use_cn:
        st.dt *= 0.2
    c = st.concen_pd[..., 0].copy()
    st.concen_pd[...] = 0.0
    st.concen_pd[9:18, 7:15, 4:9, :, 0] = c[9:18, 7:15, 4:9, :]
    st.concen_pd[..., 1] = st.concen_pd[..., 0]
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a)
    T = tr.Transport(b)
    vent = (12, 10, 5)
    for itime in range(1, 4):
        O.step_glue()
        T.step_glue()
        rc, bo = O.transport_step_subgrid(itime, True, *vent)
        bg = T.transport_step_subgrid(itime, True, *vent)
        assert rc == 0 and bg == bo
    T.sync_to_host()
    T.close()
    same(a, b)
