"""A run whose source term comes from the source builders (Suzuki column,
three pulses, one of them starting inside a step) on the device context
and on the oracle: the states are bitwise equal after every step."""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_builders_drive_the_device_run(pkg, oracle):
    tr = importlib.import_module("volcano-ash3d_amd.transport")
    srcm = importlib.import_module("volcano-ash3d_amd.source")
    st = pkg.workloads.random_state(20, 16, 14, 3, seed=91, latlon=True, use_cn=True)
    st.concen_pd[...] = 0.0
    dz = np.full(st.nz, 0.75)
    s_lb = np.concatenate([[0.0], np.cumsum(dz)])
    s_cc = s_lb[:-1] + 0.5 * dz
    kw = dict(start_time=[0.1, 0.5, 0.62], duration=[0.3, 0.1, 0.5], volume=[0.002, 0.001, 0.004],
              plume_height=[7.0, 5.0, 9.5], s_lb=s_lb, s_cc=s_cc, ds=dz, bin_mass=[0.2, 0.5, 0.3], suzuki_a=8.0)
    a, b = st.copy(), st.copy()
    O, T = oracle.Oracle(a), tr.Transport(b)
    SO, ST = srcm.Source("suzuki", **kw), srcm.Source("suzuki", **kw)
    iv, jv = 9, 7
    t, inserted = 0.0, 0
    for itime in range(1, 25):  # t = 0 .. 0.75 h: all three pulses
        for X, S in ((O, SO), (T, ST)):
            X.step_glue()
            f = S.step(X, t, st.dt, iv, jv)
            X.transport_step(itime, True)
        inserted += f is not None
        t += st.dt
        T.sync_to_host()
        assert np.array_equal(a.concen_pd, b.concen_pd), itime
    assert inserted >= 3
    T.close()
