"""Writes the gauge-file fixtures from the reference's own writers.

GaugeField::save (src/gauge.h:38-47, the raw format) and save_gauge
(src/hmc_forces.cpp:18-38, the 'SWNG'-headed format) are called through the compiled
reference (oracle/_ref/libref.so, built by oracle/Makefile from the reference's
sources) on the L = 8 field random_gauge(8, 0.6, 5), beta = 4, mass = 0.1.  Run in the
container that holds the reference:

    make -C oracle && python tests/golden/make_gauge_files.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle.ref as R  # noqa: E402
from schwinger_mg_amd.fields import random_gauge  # noqa: E402

L, SIGMA, SEED, BETA, MASS = 8, 0.6, 5, 4.0, 0.1
RAW = os.path.join(HERE, "gauge_L8_raw.bin")
SWNG = os.path.join(HERE, "gauge_L8_swng.bin")


def main():
    assert R.available(), "oracle/_ref/libref.so is not built"
    U = random_gauge(L, SIGMA, SEED)
    assert R.gauge_save_raw(L, U, RAW) == 0
    assert R.gauge_save(L, U, BETA, MASS, SWNG) == 0
    print("wrote", RAW, SWNG)


if __name__ == "__main__":
    main()
