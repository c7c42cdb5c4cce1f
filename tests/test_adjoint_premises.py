"""CPU checks of the facts the discrete adjoint rests on, on the oracle.

The adjoint of a limited DCU sweep applies J(q)^T, where J(q) is the linear
map of the limiter branches taken at q, and F(q) = J(q) q.  That needs the
sweep to be positively homogeneous of degree one: scaling q by a power of
two scales every difference, leaves every limiter ratio theta unchanged,
and is exact in floating point, so F(2q) == 2 F(q) bit for bit (away from
the EPS_THRESH skip, which the seeded cases do not straddle).  The
diffusion sweeps are linear: F(a + b) = F(a) + F(b) to rounding.
"""
import numpy as np
import pytest


def interior(st, a):
    return a[2:st.nx + 2, 2:st.ny + 2, 2:st.nz + 2]


def run(oracle, st, calls):
    a = st.copy()
    O = oracle.Oracle(a)
    for name, *args in calls:
        getattr(O, name)(*args)
    return interior(a, a.concen_pd[..., 0]).copy()


@pytest.mark.parametrize("lim", ["superbee", "minmod", "mc", "fromm", "bw", "laxwen"])
@pytest.mark.parametrize("sweep", ["advect_x", "advect_y", "advect_z"])
def test_dcu_sweeps_are_positively_homogeneous(pkg, oracle, lim, sweep):
    st = pkg.workloads.random_state(11, 9, 7, 2, seed=3, latlon=True, limiter=pkg.state.LIMITER_NAMES[lim])
    f1 = run(oracle, st, [(sweep,)])
    st2 = st.copy()
    st2.concen_pd *= 2.0
    f2 = run(oracle, st2, [(sweep,)])
    assert np.array_equal(f2, 2.0 * f1)


def test_whole_step_is_positively_homogeneous(pkg, oracle):
    st = pkg.workloads.random_state(11, 9, 7, 2, seed=4, latlon=True, periodic=True)
    f1 = run(oracle, st, [("transport_step", 1, True)])
    st2 = st.copy()
    st2.concen_pd *= 4.0
    f2 = run(oracle, st2, [("transport_step", 1, True)])
    assert np.array_equal(f2, 4.0 * f1)


@pytest.mark.parametrize("name", ["diffCN_x", "diffCN_y", "diffCN_z", "diff_x", "diff_y", "diff_z"])
def test_diffusion_sweeps_are_linear(pkg, oracle, name):
    sa = pkg.workloads.random_state(10, 9, 8, 2, seed=5, latlon=True)
    sb = pkg.workloads.random_state(10, 9, 8, 2, seed=6, latlon=True)
    sb.concen_pd[..., 0] = -0.5 * sb.concen_pd[..., 0]  # same geometry as sa, other cloud
    ss = sa.copy()
    ss.concen_pd[..., 0] += sb.concen_pd[..., 0]
    sbb = sa.copy()
    sbb.concen_pd[..., 0] = sb.concen_pd[..., 0]
    fa, fb, fs = run(oracle, sa, [(name,)]), run(oracle, sbb, [(name,)]), run(oracle, ss, [(name,)])
    scale = np.abs(fa).max() + np.abs(fb).max()
    assert np.abs(fs - (fa + fb)).max() <= 1e-13 * scale
