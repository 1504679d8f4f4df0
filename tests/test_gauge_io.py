"""Gauge field files (host-only entry points of the C-ABI, no GPU): the raw format of
GaugeField::save/load (src/gauge.h:38-57) and the 'SWNG'-headed format of
save_gauge/load_gauge (src/hmc_forces.cpp:18-62) are byte-identical to the files the
compiled reference writes, each side reads the other's files, and the error paths
(wrong L, wrong magic) refuse the file as the reference does."""
import os

import numpy as np
import pytest

import oracle.ref as R
from schwinger_mg_amd.fields import random_gauge

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "schwinger_mg_amd", "libsmg.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libsmg.so not built")


@pytest.fixture
def smg():
    import schwinger_mg_amd as m
    m.load_library()
    return m


def test_raw_round_trip_and_refusals(smg, tmp_path):
    L = 8
    U = random_gauge(L, 0.5, 3)
    p = str(tmp_path / "raw.bin")
    smg.gauge_save_raw(L, U, p)
    raw = open(p, "rb").read()
    assert len(raw) == 4 + 2 * L * L * 16
    assert int.from_bytes(raw[:4], "little") == L * L and raw[4:] == U.tobytes()
    assert np.array_equal(smg.gauge_load_raw(L, p), U)
    with pytest.raises(smg.SMGError):
        smg.gauge_load_raw(L + 2, p)   # volume mismatch


def test_swng_header_and_normalisation(smg, tmp_path):
    L = 6
    U = random_gauge(L, 0.5, 4) * 1.001   # links off the unit circle: reading renormalises
    p = str(tmp_path / "cfg.bin")
    smg.save_gauge(L, U, 2.25, 0.05, p)
    raw = open(p, "rb").read()
    assert raw[:4] == b"GNWS" and len(raw) == 32 + 2 * L * L * 16   # 0x53574E47 little-endian
    hdr = np.frombuffer(raw[8:32], dtype=np.float64)
    assert hdr[0] == 2.25 and hdr[1] == 0.05
    assert raw[32:] == U.tobytes()
    V2, h = smg.load_gauge(L, p)
    assert h.magic == 0x53574E47 and h.L == L and h.beta == 2.25 and h.mass == 0.05
    assert np.allclose(V2, U / np.abs(U), rtol=0, atol=1e-15) and np.allclose(np.abs(V2), 1.0, atol=1e-15)
    with pytest.raises(smg.SMGError):
        smg.load_gauge(L + 2, p)
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(b"XXXX" + raw[4:])
    with pytest.raises(smg.SMGError):
        smg.load_gauge(L, bad)


def test_therm_cfg_path(smg):
    # std::to_string(double) is "%f": 2.0 -> "2.000000" -> "2.00"; 10.5 -> "10.5"
    assert smg.therm_cfg_path(64, 2.0, 200) == "gauge_L64_b2.00_t200.bin"
    assert smg.therm_cfg_path(128, 10.5, 50) == "gauge_L128_b10.5_t50.bin"
    assert smg.therm_cfg_path(32, 0.125, 7) == "gauge_L32_b0.12_t7.bin"


@pytest.mark.skipif(not R.available(), reason="oracle/_ref/libref.so not built (reference sources absent)")
def test_files_byte_identical_to_the_reference(smg, tmp_path):
    L = 8   # V = 64: the reference's plaquette sum runs single-threaded (OMP_MIN_SIZE = 4096)
    U = random_gauge(L, 0.6, 5)
    a, b = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    smg.gauge_save_raw(L, U, a)
    assert R.gauge_save_raw(L, U, b) == 0
    assert open(a, "rb").read() == open(b, "rb").read()
    assert np.array_equal(R.gauge_load_raw(L, a), smg.gauge_load_raw(L, b))
    smg.save_gauge(L, U, 4.0, 0.1, a)
    assert R.gauge_save(L, U, 4.0, 0.1, b) == 0
    assert open(a, "rb").read() == open(b, "rb").read()
    Ur, hr = R.gauge_load(L, a)
    Ug, hg = smg.load_gauge(L, b)
    assert np.array_equal(Ur, Ug) and hr == (hg.beta, hg.mass, hg.avg_plaq)
    assert R.gauge_load(L + 2, a) is None


def test_reference_written_fixtures(smg, tmp_path):
    """The committed fixtures written by the reference (tests/golden/make_gauge_files.py)
    equal this library's files for the same field."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    L = 8
    U = random_gauge(L, 0.6, 5)
    a, b = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    smg.gauge_save_raw(L, U, a)
    smg.save_gauge(L, U, 4.0, 0.1, b)
    assert open(a, "rb").read() == open(os.path.join(gold, "gauge_L8_raw.bin"), "rb").read()
    assert open(b, "rb").read() == open(os.path.join(gold, "gauge_L8_swng.bin"), "rb").read()
