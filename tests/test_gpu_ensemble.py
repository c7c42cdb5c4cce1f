"""Batched ensembles (ash3d_tr_ensemble_*): several members on grids of
different sizes, each with its own dt, stepped by one batched launch per
sweep kind and kernel variant (one HIP graph per step).

* The members of the ensemble workload scaled down (odd and even nx, four
  members), a periodic lon-lat member and a member with topography, one
  with a partial IsAloft mask: after three ensemble steps every array of
  every member is bitwise what the members' own steps (step_glue +
  transport_step) give, and within the 1e-12 relative L-inf bar of the
  partitioned line solves of the oracle's dgtsv-order ones.
* Graph replays against eager batched launches, with eager work on the
  members (mass totals) between the steps: bitwise, replay after replay.
* Full-size members whose x rows take every row kernel (one wavefront per
  row in cell pairs and in single cells, segmented rows, workgroup rows):
  batched bitwise equal to the members' own steps.
* Members the ensemble cannot take are refused with EINVAL.
"""
import numpy as np
import pytest

from test_gpu_parity import ARRAYS, assert_same

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def small_members(pkg):
    W = pkg.workloads
    sts = [W.ensemble_member(m, ns=4, size_scale=0.25) for m in range(4)]
    sts.append(W.random_state(37, 21, 13, 3, seed=11, latlon=True, periodic=True))
    sts.append(W.random_state(40, 18, 11, 5, seed=12, zscaling_id=2))
    sts[1].is_aloft = np.array([1, 0, 1, 1], dtype=np.int32)
    return sts


# the members' own eager steps with the solvers an ensemble step runs: no
# Jacobi sweeps (a captured or batched step keeps the partitioned solves)
NO_JACOBI = {"jacobi_sweeps": 0}


def contexts(tr, sts, **tune):
    Ts = []
    for st in sts:
        T = tr.Transport(st)
        T.tune(line_solver="partitioned", **(NO_JACOBI | tune))
        T.transport_step(1, True)
        Ts.append(T)
    return Ts


def run_individual(tr, st, steps, **tune):
    b = st.copy()
    T = tr.Transport(b)
    T.tune(line_solver="partitioned", **(NO_JACOBI | tune))
    T.transport_step(1, True)
    for it in range(2, 2 + steps):
        T.step_glue()
        T.transport_step(it, True)
    T.sync_to_host()
    T.close()
    return b


def run_oracle(oracle, st, steps):
    a = st.copy()
    O = oracle.Oracle(a)
    O.transport_step(1, True)
    for it in range(2, 2 + steps):
        O.step_glue()
        O.transport_step(it, True)
    return a


def run_batched(tr, sts, steps, use_graph=True, between=None, streams=0, **tune):
    cs = [st.copy() for st in sts]
    Ts = contexts(tr, cs, **tune)
    E = tr.Ensemble(Ts, streams=streams)
    for it in range(2, 2 + steps):
        E.step(it, use_graph)
        if between:
            between(Ts, it)
    E.synchronize()
    for T in Ts:
        T.sync_to_host()
    info = E.info(2), E.info(3)
    E.close()
    for T in Ts:
        T.close()
    return cs, info


def test_batched_matches_members_and_oracle(pkg, oracle, tr):
    sts = small_members(pkg)
    out, info = run_batched(tr, sts, 3)
    print("launches per step (batched, members'):", info, flush=True)
    for k, st in enumerate(sts):
        print(f"member {k}: {st.nx}x{st.ny}x{st.nz}x{st.ns} dt {st.dt:.4g}", flush=True)
        assert_same(run_individual(tr, st, 3), out[k])
        assert_same(run_oracle(oracle, st, 3), out[k], tol=TOL)
    for b, m in info:
        assert b < m


@pytest.mark.parametrize("streams", [1, 2, 6])
def test_stream_groups(pkg, tr, streams):
    """One, two and six stream groups (graph branches) give the bits of
    the members' own steps."""
    sts = small_members(pkg)
    out, _ = run_batched(tr, sts, 2, streams=streams)
    for k, st in enumerate(sts):
        assert_same(run_individual(tr, st, 2), out[k])


@pytest.mark.parametrize("fused", [0, 1])
def test_advection_modes(pkg, oracle, tr, fused):
    """Members with the split sweeps and with the fused second horizontal +
    vertical pass (the batched fused kernel): bitwise the members' own
    steps in the same mode, 1e-12 of the oracle."""
    sts = small_members(pkg)
    out, _ = run_batched(tr, sts, 2, adv_fused=fused)
    for k, st in enumerate(sts):
        assert_same(run_individual(tr, st, 2, adv_fused=fused), out[k])
        assert_same(run_oracle(oracle, st, 2), out[k], tol=TOL)


def test_glue_and_source_between_steps(pkg, tr):
    """The reference's loop body between steps on the members -- step glue,
    then the vent-column source -- followed by the batched step: bitwise
    the members' own glue, source and step."""
    sts = small_members(pkg)

    def flux(st, it):
        rng = np.random.default_rng(100 * it + st.nx)
        return np.asfortranarray(rng.random((st.nz + 1, st.ns)) * 50.0)

    def vent(st):
        return max(1, st.nx // 3), max(1, st.ny // 2)

    own = []
    for st in sts:
        b = st.copy()
        T = tr.Transport(b)
        T.tune(line_solver="partitioned", **NO_JACOBI)
        T.transport_step(1, True)
        for it in range(2, 5):
            T.step_glue()
            T.source_column_add(*vent(st), flux(st, it))
            T.transport_step(it, True)
        T.sync_to_host()
        T.close()
        own.append(b)
    cs = [st.copy() for st in sts]
    Ts = contexts(tr, cs)
    E = tr.Ensemble(Ts)
    for it in range(2, 5):
        for T, st in zip(Ts, sts):
            T.step_glue()
            T.source_column_add(*vent(st), flux(st, it))
        E.step(it)
    E.synchronize()
    for T in Ts:
        T.sync_to_host()
    E.close()
    for T in Ts:
        T.close()
    for a, b in zip(own, cs):
        assert_same(a, b)


def test_graph_replays_match_eager(pkg, tr):
    sts = small_members(pkg)
    seen = {True: [], False: []}

    def totals(flag):
        def f(Ts, it):
            seen[flag].append([np.concatenate(T.mass_totals()) for T in Ts])
        return f

    g, _ = run_batched(tr, sts, 4, use_graph=True, between=totals(True))
    e, _ = run_batched(tr, sts, 4, use_graph=False, between=totals(False))
    for a, b in zip(g, e):
        assert_same(a, b)
    for ra, rb in zip(seen[True], seen[False]):
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)


# members 1 (394 cells, even: segmented rows), 2 (357, odd: single-cell
# wavefront rows), 4 and 5 (122, 126: pair rows of one and two chunks) and
# 35 (381, odd beyond the wavefront kernel: workgroup rows)
FULL = [1, 2, 4, 5, 35]


def test_full_size_members_match_their_own_steps(pkg, tr):
    sts = [pkg.workloads.ensemble_member(m, ns=3) for m in FULL]
    out, info = run_batched(tr, sts, 2)
    print("launches per step (batched, members'):", info, flush=True)
    for k, st in enumerate(sts):
        print(f"member {FULL[k]}: {st.nx}x{st.ny}x{st.nz}x{st.ns}", flush=True)
        assert_same(run_individual(tr, st, 2), out[k])


def test_refuses_members_it_cannot_batch(pkg, tr):
    W = pkg.workloads
    st = W.random_state(20, 12, 9, 2, seed=5, latlon=True)
    T = tr.Transport(st.copy())
    T.tune(line_solver="partitioned")
    with pytest.raises(tr.Ash3dError) as ex:  # not stepped yet
        tr.Ensemble([T])
    assert ex.value.code == 1
    T.transport_step(1, True)
    T.tune(line_solver="exact")
    with pytest.raises(tr.Ash3dError) as ex:  # dgtsv-order solves
        tr.Ensemble([T])
    assert ex.value.code == 1
    T.tune(line_solver="partitioned")
    with pytest.raises(tr.Ash3dError):  # the same member twice
        tr.Ensemble([T, T])
    with pytest.raises(tr.Ash3dError):  # more than 16 stream groups
        tr.Ensemble([T], streams=17)
    E = tr.Ensemble([T])
    E.step(2)
    E.synchronize()
    # new diffusivities after the ensemble was built: the line factors are
    # rebuilt and checked again -- negative ones make lines that the
    # partitioned solves cannot take, and the step says so
    T.st.kx[...] = -5.0
    T.sync_from_host(tr.SYNC_DIFFK)
    with pytest.raises(tr.Ash3dError) as ex:
        E.step(3)
    assert ex.value.code == 1
    E.close()
    T.close()
