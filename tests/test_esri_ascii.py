"""ESRI ASCII products (write_2D_ASCII, src/write_ASCII.f90:264-382) and the
oracle's product calculators (src/Output_Vars.f90:855-893, 1026-1137).

The writer is pinned byte for byte against output files of the reference's
own regression tests (tests/golden/esri/, copied by copy_fixtures.py): each
file is parsed back to its grid, header and masked cells, written again by
the library, and must come out identical -- header formats (i5, f15.3,
a6), the km-to-m and longitude conventions, real(4) values in 10f18.6
records from the northern row down, the blank record after each row.
The writer is host code: no device is used.
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "esri", "*.dat")))


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


def parse(path):
    lines = open(path).read().split("\n")
    nx = int(lines[0].split()[1])
    ny = int(lines[1].split()[1])
    xll = float(lines[2].split()[1])
    yll = float(lines[3].split()[1])
    dx, dy = (float(v) for v in lines[4].split()[1:3])
    fill = lines[5].split()[1]
    nums = []
    for line in lines[6:]:
        if line.strip():
            nums += [float(line[i:i + 18]) for i in range(0, len(line), 18)]
    rows = np.array(nums).reshape(ny, nx)  # northern row first
    return nx, ny, xll, yll, dx, dy, fill, rows[::-1].T.copy()  # (nx, ny), j = 1 south


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_writer_reproduces_reference_files(tr, tmp_path, path):
    nx, ny, xll, yll, dx, dy, fill, field = parse(path)
    latlon = "t04" in os.path.basename(path)  # test 4 is the lon-lat case
    if not latlon:  # the reference writes km inputs in metres
        xll, yll, dx, dy = xll / 1e3, yll / 1e3, dx / 1e3, dy / 1e3
    mask = (field != float(fill)).astype(np.int32)
    field = np.where(mask == 1, field, 12345.0)  # masked cells must not be read
    out = str(tmp_path / "out.dat")
    tr.write_esri_ascii(out, field, mask, fill, latlon, xll, yll, dx, dy)
    assert open(out).read() == open(path).read()


def test_writer_conventions(tr, tmp_path):
    """lonLL > 180 is shifted by -360; values that do not fit f18.6 print as
    asterisks; values go through real(4); rows run north to south."""
    f = np.zeros((3, 2))
    f[0, 0] = 1.0 / 3.0          # real(4): 0.333333
    f[1, 0] = 2.0e11             # does not fit f18.6
    f[2, 1] = -7.0
    out = str(tmp_path / "c.dat")
    tr.write_esri_ascii(out, f, None, "-9999.", True, 200.0, -10.0, 0.25, 0.5)
    lines = open(out).read().split("\n")
    assert lines[2] == "XLLCORNER " + f"{-160.0:15.3f}"
    assert lines[3] == "YLLCORNER " + f"{-10.0:15.3f}"
    assert lines[4] == "CELLSIZE " + f"{0.25:15.3f}" + f"{0.5:15.3f}"
    assert lines[5] == "NODATA_VALUE -9999."
    assert lines[6] == f"{0.0:18.6f}{0.0:18.6f}{-7.0:18.6f}"  # j = 2 first
    assert lines[7] == "  "
    assert lines[8] == f"{float(np.float32(1 / 3)):18.6f}" + "*" * 18 + f"{0.0:18.6f}"
    assert lines[9] == "  " and lines[10] == ""


def test_file_names(tr):
    assert tr.esri_ascii_name("CloudLoad_          ", 4.0) == "CloudLoad_004.00hrs.dat"
    assert tr.esri_ascii_name("CloudHeight_        ", 16.0) == "CloudHeight_016.00hrs.dat"
    assert tr.esri_ascii_name("DepositFile_        ", 123.5) == "DepositFile_123.50hrs.dat"
    assert tr.esri_ascii_name("DepositFile_        ", 8.0, True) == "DepositFile_____final.dat"
    assert tr.esri_ascii_name("CloudArrivalTime    ", 3.0) == "CloudArrivalTime.dat"
    assert tr.esri_ascii_name("outvar", 3.0) == "outvar.dat"
    # the fixtures' own names
    for p in FIXTURES:
        name = os.path.basename(p)[4:]
        root = name.split("_")[0] + "_"
        if "ArrivalTime" in name:
            assert tr.esri_ascii_name(name[:-4], 0.0) == name
        elif "final" in name:
            assert tr.esri_ascii_name(root, 0.0, True) == name
        else:
            assert tr.esri_ascii_name(root, float(name[len(root):len(root) + 6])) == name


def test_oracle_products_hand_values(pkg, oracle):
    """One column with a hand-chosen profile: load = sum_k TC_k dz_k / 1e3,
    cloud top = top of the highest level with TC >= thresh, bottom = bottom
    of the lowest level with TC <= thresh, peak in mg/m3, masks and areas;
    deposit thickness = sum(DG) dz_1 / 1e6 / rho * 1e3."""
    st = pkg.workloads.random_state(4, 3, 6, 2, seed=3)
    st.concen_pd[...] = 0.0
    st.deposit_granularity[...] = 0.0
    dz = np.array([0.5, 0.5, 1.0, 1.0, 2.0, 2.0])
    zcc = np.cumsum(dz) - 0.5 * dz
    # column (2, 2): total concentrations per level (two bins), ts1
    tc = np.array([0.0, 5e-4, 400.0, 1000.0, 2e-3, 0.0])
    for k in range(6):
        st.concen_pd[3, 3, k + 2, 0, 1] = 0.25 * tc[k]
        st.concen_pd[3, 3, k + 2, 1, 1] = 0.75 * tc[k]
    # column (4, 1): a thin layer below the load threshold
    st.concen_pd[5, 2, 4, 0, 1] = 1e-2
    st.deposit_granularity[0, 0, :] = [3.0e5, 1.0e5]
    st.bounds = (1, 4, 1, 3, 1, 6)
    r = oracle.Oracle(st).products(dz, zcc, cloudcon_thresh=1e-3, cloudload_thresh=0.2, depo_thresh=1e-2,
                                   deposit_density=800.0)
    load = float(np.sum(tc * dz) / 1e3)
    assert r["cloud_load"][1, 1] == pytest.approx(load, rel=1e-15)
    assert r["mask_cloud"][1, 1] == 1 and r["mask_cloud"].sum() == 1
    assert r["max_height"][1, 1] == pytest.approx(zcc[4] + 0.5 * dz[4])  # level 5 (2e-3 >= 1e-3)
    assert r["min_height"][1, 1] == pytest.approx(0.0)  # level 1 bottom (TC = 0 <= thresh)
    assert r["max_concentration"][1, 1] == pytest.approx(1000.0 * 1e6 / 1e9)
    assert r["cloud_load"][3, 0] == 0.0 and r["max_height"][3, 0] == 0.0  # below CLOUDLOAD_THRESH
    sig = st.sigma_nz_pd[3, 3, 2]
    assert r["cloud_area"] == sig
    th = 4.0e5 * dz[0] / 1e6 / 800.0 * 1e3
    assert r["deposit_thickness"][0, 0] == pytest.approx(th, rel=1e-15)
    assert r["mask_deposit"][0, 0] == 1 and r["mask_deposit"].sum() == 1
    assert r["deposit_area"] == st.sigma_nz_pd[2, 2, 2]
    # outside the box ConcentrationCalculator leaves MinHeight at 100
    st.bounds = (1, 2, 1, 3, 1, 6)
    r = oracle.Oracle(st).products(dz, zcc)
    assert r["min_height"][3, 0] == 100.0 and r["cloud_load"][1, 1] > 0
