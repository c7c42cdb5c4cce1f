"""x advection rows against the oracle, bitwise, with both row kernels:
one wavefront per row (the row in registers, as 16-byte cell pairs for
even nx and single cells for odd nx, the neighbours by DPP rotations
across chunks of 128 or 64 cells) and one workgroup per row (the row
staged through LDS).  Row lengths sit at and around the chunk edges
(nx + 4 = 64, 128, 256, 384 and the pair/single-cell limits).  Longer
even rows (nx > 380, > 252 with topography) take several wavefronts per
row, a segment each with the halos exchanged through LDS, up to 8
segments; odd ones, longer ones and more than 64 bins take the workgroup
kernel whatever the setting.  The adjoint sweep has the same three even-row
kernels (segments of at most 248 cells) and is held to the workgroup
kernel's bits and to the dot-product identity with the oracle's forward
sweep."""
import numpy as np
import pytest

from test_gpu_parity import assert_same, run_both

pytestmark = pytest.mark.gpu

NX = [2, 12, 124, 126, 252, 254, 380, 382, 508, 510,
      3, 13, 59, 61, 123, 125, 251, 253, 379, 381]
VARIANTS = {
    "cartesian": dict(latlon=False, periodic=False, zscaling_id=0),
    "lonlat_periodic": dict(latlon=True, periodic=True, zscaling_id=0),
    "topography": dict(latlon=False, periodic=False, zscaling_id=2),
    "subgrid_aloft": dict(latlon=True, periodic=False, zscaling_id=0),
}


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def advect_x_both(pkg, oracle, tr, st, xrows):
    a, b = st.copy(), st.copy()
    O, T = oracle.Oracle(a), tr.Transport(b)
    T.tune(adv_x_rows=xrows)
    O.advect_x()
    T.advect_x()
    T.sync_to_host()
    T.close()
    return a, b


@pytest.mark.parametrize("xrows", [1, 0])
@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("nx", NX)
def test_x_rows_kernels(pkg, oracle, tr, nx, variant, xrows):
    ns = 5
    st = pkg.workloads.random_state(nx, 5, 4, ns, seed=nx + 31 * len(variant), **VARIANTS[variant])
    if variant == "subgrid_aloft":
        lo = min(3, nx)
        st.bounds = (lo, max(lo, nx - 2), 2, 4, 2, 3)
        st.is_aloft = np.array([1, 0, 1, 1, 0], dtype=np.int32)
    a, b = advect_x_both(pkg, oracle, tr, st, xrows)
    assert_same(a, b)


# segmented rows: 2 and 3 segments, the global grid's 1444-cell rows (4
# segments; 6 with topography), a segment at its 380-cell cap, one cell
# past it (5 segments), 8 segments and the first row past 8 (workgroup
# kernel)
NX_SEG = [382, 762, 1440, 1516, 1518, 3036, 3038]


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("nx", NX_SEG)
def test_x_segmented_rows(pkg, oracle, tr, nx, variant):
    ns = 5
    st = pkg.workloads.random_state(nx, 3, 3, ns, seed=nx + 17 * len(variant), **VARIANTS[variant])
    if variant == "subgrid_aloft":
        st.bounds = (3, nx - 2, 2, 3, 2, 3)
        st.is_aloft = np.array([1, 0, 1, 1, 0], dtype=np.int32)
    a, b = advect_x_both(pkg, oracle, tr, st, 1)
    assert_same(a, b)


@pytest.mark.parametrize("ns", [1, 2, 20, 64])
def test_x_segmented_rows_bin_counts(pkg, oracle, tr, ns):
    st = pkg.workloads.random_state(1440, 2, 3, ns, seed=100 + ns, latlon=True, periodic=True)
    st.is_aloft = np.ones(ns, dtype=np.int32)
    st.is_aloft[1::4] = 0 if ns > 2 else 1
    a, b = advect_x_both(pkg, oracle, tr, st, 1)
    assert_same(a, b)


@pytest.mark.parametrize("ns", [1, 2, 15, 64, 65])
def test_x_rows_bin_counts(pkg, oracle, tr, ns):
    """Odd and even bin counts (the wavefront kernel takes bins in pairs),
    the last bin count it takes (64: one outflow lane per bin) and the
    first it leaves to the workgroup kernel."""
    st = pkg.workloads.random_state(126, 3, 3, ns, seed=ns, latlon=True)
    st.is_aloft = np.ones(ns, dtype=np.int32)
    st.is_aloft[::3] = 0 if ns > 2 else 1
    a, b = advect_x_both(pkg, oracle, tr, st, 1)
    assert_same(a, b)


def test_x_rows_in_a_step_sequence(pkg, oracle, tr):
    """Two full steps with diffusion on an even lon-lat grid (x advection
    first in one step, second in the other), bitwise."""
    st = pkg.workloads.random_state(60, 22, 9, 7, seed=77, latlon=True, zscaling_id=2)
    st.concen_pd[..., 1] = st.concen_pd[..., 0]
    calls = []
    for itime in (1, 2):
        calls += [("step_glue",), ("transport_step", itime, True)]
    a, b = run_both(pkg, oracle, tr, st, calls)
    assert_same(a, b)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("nx", NX)
def test_adjoint_x_rows_kernels(pkg, oracle, tr, nx, variant):
    """The adjoint x sweep: the wavefront kernel gives the bits of the
    workgroup kernel, and both satisfy <F(q), w> = <q, J(q)^T w> with the
    oracle's forward sweep F (the identity the adjoint tests pin)."""
    import test_gpu_adjoint as ta
    st = pkg.workloads.random_state(nx, 5, 4, 5, seed=nx + 7 * len(variant), **VARIANTS[variant])
    if variant == "subgrid_aloft":
        lo = min(3, nx)
        st.bounds = (lo, max(lo, nx - 2), 2, 4, 2, 3)
        st.is_aloft = np.array([1, 0, 1, 1, 0], dtype=np.int32)
    Fq = ta.forward(oracle, st, [("advect_x",)])
    lam = ta.rand_lambda(st, nx)
    g1 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=1)
    g0 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=0)
    assert np.array_equal(g1, g0)
    ta.check_dot(st, Fq, lam, g1)


# adjoint rows in segments of one wavefront (two chunks of 128 cells, at
# most 248 cells a segment): 2 segments, the global grid's 1444-cell rows
# (6), 8 segments at the cap and the first row past it (workgroup kernel)
NX_ADJ_SEG = [382, 762, 1440, 1980, 1982]


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("nx", NX_ADJ_SEG)
def test_adjoint_x_segmented_rows(pkg, oracle, tr, nx, variant):
    """Segmented adjoint rows give the bits of the workgroup kernel and
    satisfy <F(q), w> = <q, J(q)^T w> with the oracle's forward sweep."""
    import test_gpu_adjoint as ta
    st = pkg.workloads.random_state(nx, 3, 3, 5, seed=nx + 11 * len(variant), **VARIANTS[variant])
    if variant == "subgrid_aloft":
        st.bounds = (3, nx - 2, 2, 3, 2, 3)
        st.is_aloft = np.array([1, 0, 1, 1, 0], dtype=np.int32)
    Fq = ta.forward(oracle, st, [("advect_x",)])
    lam = ta.rand_lambda(st, nx + 1)
    g1 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=1)
    g0 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=0)
    assert np.array_equal(g1, g0)
    ta.check_dot(st, Fq, lam, g1)


@pytest.mark.parametrize("ns", [1, 2, 7])
def test_adjoint_x_segmented_rows_bin_counts(pkg, oracle, tr, ns):
    import test_gpu_adjoint as ta
    st = pkg.workloads.random_state(1440, 2, 3, ns, seed=200 + ns, latlon=True, periodic=True)
    st.is_aloft = np.ones(ns, dtype=np.int32)
    st.is_aloft[1::3] = 0 if ns > 2 else 1
    Fq = ta.forward(oracle, st, [("advect_x",)])
    lam = ta.rand_lambda(st, ns)
    g1 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=1)
    g0 = ta.adjoint(tr, st, lam, "adj_advect", 0, adv_x_rows=0)
    assert np.array_equal(g1, g0)
    ta.check_dot(st, Fq, lam, g1)
