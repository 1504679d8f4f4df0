"""Gauge files through the device: the reference-written L = 8 fixtures (raw format of
GaugeField::save, src/gauge.h:38-57, and the 'SWNG'-headed format of save_gauge,
src/hmc_forces.cpp:18-62; written by tests/golden/make_gauge_files.py) are read, the
field is uploaded into an operator handle, evolved by nothing, read back from the
device and written again: the files are byte-identical to the reference's."""
import os

import numpy as np
import pytest

from helpers import banner
from schwinger_mg_amd.fields import random_gauge

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
L, BETA, MASS = 8, 4.0, 0.1


def test_device_field_written_byte_identical_to_reference(smg, tmp_path):
    banner("gauge files through the device L=8")
    raw_ref = open(os.path.join(GOLD, "gauge_L8_raw.bin"), "rb").read()
    swng_ref = open(os.path.join(GOLD, "gauge_L8_swng.bin"), "rb").read()
    U = smg.gauge_load_raw(L, os.path.join(GOLD, "gauge_L8_raw.bin"))
    assert np.array_equal(U, random_gauge(L, 0.6, 5))
    U2, hdr = smg.load_gauge(L, os.path.join(GOLD, "gauge_L8_swng.bin"))
    assert hdr.L == L and hdr.beta == BETA and hdr.mass == MASS
    D = smg.DiracOp(L, U, MASS, 1.0, 1.0, 0.0)
    Ud = smg.gauge_of(D)        # the device-held field, read back
    assert np.array_equal(Ud, U)
    assert abs(smg.avg_plaq(L, Ud) - hdr.avg_plaq) < 1e-15
    a, b = str(tmp_path / "raw.bin"), str(tmp_path / "swng.bin")
    smg.gauge_save_raw(L, Ud, a)
    smg.save_gauge(L, Ud, BETA, MASS, b)
    assert open(a, "rb").read() == raw_ref
    assert open(b, "rb").read() == swng_ref
    # a field read from the reference's file, uploaded and applied, equals the field
    # the reference wrote (the header format renormalises |U| to 1 on reading)
    assert np.abs(U2 - U).max() < 1e-15
