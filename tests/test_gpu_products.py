"""Output products on the device (ash3d_tr_products) against the oracle's
ConcentrationCalculator and AshThicknessCalculator, and the ESRI ASCII
files written from them.

The states come from the same steps on both sides (dgtsv-order solves, so
the states are bitwise equal), hence the products must be bitwise equal:
the device sums bins and levels in the reference's order.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("cloud_load", "max_concentration", "max_height", "min_height", "deposit_thickness", "mask_cloud",
        "mask_deposit")


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def stepped(pkg, oracle, tr, nsteps=6):
    st = pkg.workloads.random_state(30, 22, 16, 4, seed=81, latlon=True, use_cn=True)
    st.concen_pd[...] = 0.0
    st.concen_pd[..., 1] = 0.0
    flux = np.zeros((st.nz + 1, st.ns), order="F")
    flux[0:8, :] = 4.0e3 * (1.0 + np.arange(st.ns))[None, :]
    a, b = st.copy(), st.copy()
    O, T = oracle.Oracle(a), tr.Transport(b)
    for itime in range(1, nsteps + 1):
        for X in (O, T):
            X.step_glue()
            X.source_column_add(12, 10, flux)
            X.transport_step(itime, True)
    return st, O, T


@pytest.mark.parametrize("bounds", [None, (5, 20, 3, 18)])
def test_products_match_oracle(pkg, oracle, tr, bounds):
    st, O, T = stepped(pkg, oracle, tr)
    dz = np.full(st.nz, 0.4) + 0.05 * np.arange(st.nz)
    zcc = np.cumsum(dz) - 0.5 * dz
    kw = dict(cloudcon_thresh=1e-3, cloudload_thresh=1e-3, depo_thresh=1e-9, deposit_density=1000.0)
    got = T.products(dz, zcc, bounds=bounds, **kw)
    if bounds is not None:
        O.st.bounds = (*bounds, 1, st.nz)
    ref = O.products(dz, zcc, **kw)
    assert ref["mask_cloud"].sum() >= 5 and ref["mask_deposit"].sum() >= 3  # the case has a cloud and a deposit
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert got["cloud_area"] == ref["cloud_area"] and got["deposit_area"] == ref["deposit_area"]
    T.close()


def test_products_after_glue_read_zero_ts1(pkg, oracle, tr):
    """After the step glue ts1 is zero: ConcentrationCalculator sees no ash."""
    st, O, T = stepped(pkg, oracle, tr, nsteps=2)
    for X in (O, T):
        X.step_glue()
    dz = np.full(st.nz, 0.5)
    zcc = np.cumsum(dz) - 0.25
    got = T.products(dz, zcc, cloudload_thresh=0.0)  # every column is "cloud" at threshold 0
    ref = O.products(dz, zcc, cloudload_thresh=0.0)
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    assert got["cloud_load"].max() == 0.0 and got["mask_cloud"].min() == 1
    T.close()


def test_esri_files_from_device_products(pkg, oracle, tr, tmp_path):
    st, O, T = stepped(pkg, oracle, tr)
    dz = np.full(st.nz, 0.5)
    zcc = np.cumsum(dz) - 0.25
    got = T.products(dz, zcc, cloudload_thresh=1e-3, depo_thresh=1e-9)
    ref = O.products(dz, zcc, cloudload_thresh=1e-3, depo_thresh=1e-9)
    for field, mask, root in (("cloud_load", None, "CloudLoad_          "),
                              ("max_height", "mask_cloud", "CloudHeight_        "),
                              ("deposit_thickness", None, "DepositFile_        ")):
        name = tr.esri_ascii_name(root, 4.0)
        files = []
        for tag, r in (("gpu", got), ("orc", ref)):
            path = str(tmp_path / f"{tag}_{name}")
            tr.write_esri_ascii(path, r[field], None if mask is None else r[mask], "-9999.", True,
                                190.0, 40.0, 0.5, 0.5)
            files.append(open(path).read())
        assert files[0] == files[1], name
    T.close()
