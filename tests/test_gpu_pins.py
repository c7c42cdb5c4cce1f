"""The hand-derived advection rows of tests/test_oracle_pins.py through the
HIP library: the same worked-out values, bitwise."""
import importlib

import pytest

import test_oracle_pins as pins

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tr(pkg):
    t = importlib.import_module("volcano-ash3d_amd.transport")
    t.load()
    return t


@pytest.mark.parametrize("name", ["test_fss_uses_kappa_of_cell_I", "test_eps_skip_at_and_above_threshold",
                                  "test_theta_clamp_both_signs", "test_topography_terms",
                                  "test_top_outflow_index_with_kmin_above_one"])
def test_pin_rows_on_device(tr, name):
    getattr(pins, name)((tr, "gpu"))
