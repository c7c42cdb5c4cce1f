"""Captured steps (HIP graphs) against eager steps: bitwise, over several
replays of both parities, with eager work and host syncs between replays
(each replay is checked, not only the last)."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr(pkg):
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


@pytest.mark.parametrize("case", [dict(latlon=True, periodic=True, use_cn=True),
                                  dict(latlon=False, periodic=False, use_cn=False),
                                  dict(latlon=True, periodic=False, use_cn=True, zscaling_id=2)])
def test_graph_replays_match_eager(pkg, tr, case):
    st = pkg.workloads.random_state(21, 17, 10, 4, seed=31, **case)
    if not case["use_cn"]:
        st.dt *= 0.2
    st.concen_pd[..., 1] = st.concen_pd[..., 0]
    a, b = st.copy(), st.copy()
    E = tr.Transport(a)
    G = tr.Transport(b)
    for T in (E, G):
        T.step_glue()
        T.transport_step(1, True)
    G.graph_capture(2, True)
    G.graph_capture(3, True)
    for itime in range(2, 9):
        E.step_glue()
        E.transport_step(itime, True)
        G.graph_launch(itime)
        # eager work and a host sync between replays
        ma = E.mass_totals()
        mb = G.mass_totals()
        assert all(np.array_equal(x, y) for x, y in zip(ma, mb))
        E.sync_to_host()
        G.sync_to_host()
        assert np.array_equal(a.concen_pd, b.concen_pd), itime
        for name in ("outflow_xz1_pd", "outflow_yz2_pd", "outflow_xy1_pd", "deposit_granularity"):
            assert np.array_equal(getattr(a, name), getattr(b, name)), (itime, name)
    # an upload makes ts1 the host's copy: the captured glue no longer applies
    G.sync_from_host(tr.SYNC_CONCEN)
    with pytest.raises(tr.Ash3dError, match="state changed"):
        G.graph_launch(9)
    E.close()
    G.close()


def test_graph_replay_after_factor_changes(pkg, tr):
    """The captured line solves read cached row and LU factors: an eager
    step at another dt (which refactors the lines) and a diffusivity upload
    between replays must not leak into the replays."""
    st = pkg.workloads.random_state(21, 17, 10, 4, seed=33, latlon=True, periodic=False, use_cn=True)
    st.concen_pd[..., 1] = st.concen_pd[..., 0]
    a, b = st.copy(), st.copy()
    E, G = tr.Transport(a), tr.Transport(b)
    dt0 = st.dt
    for T in (E, G):
        T.step_glue()
        T.transport_step(1, True)
    G.graph_capture(2, True)
    G.graph_capture(3, True)

    def check(itime):
        E.sync_to_host()
        G.sync_to_host()
        assert np.array_equal(a.concen_pd, b.concen_pd), itime

    for X, T in ((a, E), (b, G)):  # eager step at half the dt
        X.dt = 0.5 * dt0
        T.step_glue()
        T.transport_step(2, True)
        X.dt = dt0
    check(2)
    E.step_glue()
    E.transport_step(3, True)
    G.graph_launch(3)
    check(3)
    for X, T in ((a, E), (b, G)):  # new diffusivities
        X.kx *= 1.5
        X.kz *= 0.5
        T.sync_from_host(tr.SYNC_DIFFK)
    for itime in (4, 5):
        E.step_glue()
        E.transport_step(itime, True)
        G.graph_launch(itime)
        check(itime)
    E.close()
    G.close()


def test_graph_capture_needs_an_eager_step(pkg, tr):
    st = pkg.workloads.random_state(9, 8, 6, 2, seed=2)
    T = tr.Transport(st.copy())
    with pytest.raises(tr.Ash3dError, match="eager step"):
        T.graph_capture(1, True)
    T.step_glue()
    T.transport_step(1, True)
    with pytest.raises(tr.Ash3dError, match="no graph"):
        T.graph_launch(2)
    T.close()
