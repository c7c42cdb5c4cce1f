"""The fused advection of the step with diffusion (ASH3D_TUNE_ADV_FUSED,
the default): the first horizontal sweep into the ts1 slot, then the
second horizontal sweep and the vertical sweep in one pass from ts1 into
ts0 (tiles of 64 x 8 columns streamed upward through LDS planes).  Whole
steps (both parities, so both x-then-y and y-then-x) against the oracle
and against the split sweeps, bitwise, on grids whose sizes sit around
the tile edges (nx around 64, ny around 8), periodic lon-lat grids (the
x halo wrapped), topography, every limiter, partial IsAloft masks and bin
counts from 1 to 9."""
import numpy as np
import pytest

from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def steps(pkg, oracle, tr, st, n_steps=2, fused=1, line_solver="exact", with_oracle=True):
    a, b = st.copy(), st.copy()
    O = oracle.Oracle(a) if with_oracle else None
    T = tr.Transport(b)
    T.tune(adv_fused=fused, line_solver=line_solver)
    for it in range(1, n_steps + 1):
        if it > 1:
            if O:
                O.step_glue()
            T.step_glue()
        if O:
            O.transport_step(it, True)
        T.transport_step(it, True)
    T.sync_to_host()
    T.close()
    return a, b


SHAPES = [(5, 3, 4), (63, 7, 6), (64, 8, 5), (65, 9, 7), (130, 17, 5), (37, 22, 9)]
CASES = {
    "cartesian": dict(latlon=False, periodic=False, zscaling_id=0),
    "lonlat_periodic": dict(latlon=True, periodic=True, zscaling_id=0),
    "topography": dict(latlon=False, periodic=False, zscaling_id=2),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("shape", SHAPES)
def test_fused_steps_match_oracle(pkg, oracle, tr, shape, case):
    nx, ny, nz = shape
    ns = 1 + (nx + ny) % 4
    st = pkg.workloads.random_state(nx, ny, nz, ns, seed=nx * 7 + ny, **CASES[case])
    print(f"shape {shape} ns {ns} case {case}", flush=True)
    a, b = steps(pkg, oracle, tr, st)
    assert_same(a, b)


@pytest.mark.parametrize("lim", [0, 1, 2, 3, 4, 6])
def test_fused_limiters(pkg, oracle, tr, lim):
    st = pkg.workloads.random_state(66, 10, 6, 2, seed=40 + lim, latlon=True, periodic=lim % 2 == 1, limiter=lim)
    a, b = steps(pkg, oracle, tr, st)
    assert_same(a, b)


def test_fused_aloft_mask_and_bins(pkg, oracle, tr):
    st = pkg.workloads.random_state(70, 12, 8, 9, seed=91, latlon=True, zscaling_id=2)
    st.is_aloft = np.array([1, 0, 1, 1, 0, 0, 1, 0, 1], dtype=np.int32)
    a, b = steps(pkg, oracle, tr, st, n_steps=3)
    assert_same(a, b)


@pytest.mark.parametrize("case", list(CASES))
def test_fused_equals_split(pkg, oracle, tr, case):
    """The fused and split advection give the same bits, with the
    partitioned line solves too, over four steps."""
    st = pkg.workloads.random_state(140, 20, 11, 3, seed=5, **CASES[case])
    _, f = steps(pkg, oracle, tr, st, 4, fused=1, line_solver="partitioned", with_oracle=False)
    _, s = steps(pkg, oracle, tr, st, 4, fused=0, line_solver="partitioned", with_oracle=False)
    assert_same(s, f)
