"""Copy the reference's own ESRI ASCII products used as writer fixtures.

These are output files that the reference's regression tests hold
(tests/test_0*/output*/, written by write_2D_ASCII, src/write_ASCII.f90:
264-382): a projected grid (test 1: cloud load, cloud height with masked
cells, cloud arrival time) and a lon-lat grid (test 4: final deposit,
cloud concentration).  Data, not source: run once where the reference is
checked out; the copies are committed next to this script.
"""
import os
import shutil

REF = "/root/reference/tests"
FILES = {
    "t01_CloudLoad_004.00hrs.dat": "test_01/output/CloudLoad_004.00hrs.dat",
    "t01_CloudHeight_002.00hrs.dat": "test_01/output/CloudHeight_002.00hrs.dat",
    "t01_CloudArrivalTime.dat": "test_01/output/CloudArrivalTime.dat",
    "t04_DepositFile_____final.dat": "test_04/output2/DepositFile_____final.dat",
    "t04_CloudConcentration_008.00hrs.dat": "test_04/output2/CloudConcentration_008.00hrs.dat",
}

if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    for dst, src in FILES.items():
        shutil.copyfile(os.path.join(REF, src), os.path.join(here, dst))
        print(src, "->", dst)
