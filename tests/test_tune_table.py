"""The autotuner's table (host memory, no GPU): choices set by hand are read back, the
text file round-trips, and a malformed file or an impossible choice is refused without
changing the table."""
import os

import pytest

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "schwinger_mg_amd", "libsmg.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libsmg.so not built")


@pytest.fixture
def smg():
    import schwinger_mg_amd as m
    m.load_library()
    m.tune_clear()
    yield m
    m.tune_clear()


def test_set_get_save_load(smg, tmp_path):
    a, b = [128, 128, 32, 32, 2], [512, 32, 4, 0, 0]
    assert smg.tune_get(smg.TUNE_COARSE_BF16, a) == -1
    smg.tune_set(smg.TUNE_COARSE_BF16, a, 1 + 2 * 3)
    smg.tune_set(smg.TUNE_DSLASH_DOTS, b, 2)
    assert smg.tune_get(smg.TUNE_COARSE_BF16, a) == 7 and smg.tune_get(smg.TUNE_DSLASH_DOTS, b) == 2
    p = str(tmp_path / "tune.txt")
    smg.tune_save(p)
    lines = [l.split() for l in open(p) if not l.startswith("#")]
    assert sorted(lines) == sorted([["0"] + [str(w) for w in a] + ["7"], ["1"] + [str(w) for w in b] + ["2"]])
    smg.tune_clear()
    assert smg.tune_get(smg.TUNE_COARSE_BF16, a) == -1
    smg.tune_load(p)
    assert smg.tune_get(smg.TUNE_COARSE_BF16, a) == 7 and smg.tune_get(smg.TUNE_DSLASH_DOTS, b) == 2


def test_refusals(smg, tmp_path):
    with pytest.raises(smg.SMGError):
        smg.tune_set(smg.TUNE_DSLASH_DOTS, [512, 32, 4, 0, 0], 5)      # no sixth grid
    with pytest.raises(smg.SMGError):
        smg.tune_set(7, [1, 2, 3, 4, 5], 0)                              # no such family
    smg.tune_set(smg.TUNE_TRANSFER, [512, 512, 128 * 2 + 32, 32, 1], 2)       # matrix cores
    with pytest.raises(smg.SMGError):
        smg.tune_set(smg.TUNE_TRANSFER, [512, 512, 128 * 2 + 32, 32, 1], 3)   # choices 0, 1, 2
    with pytest.raises(smg.SMGError):
        smg.tune_set(smg.TUNE_CG_UPDATE, [131072, 32, 0, 0, 0], 5)         # no sixth grid
    smg.tune_set(smg.TUNE_DDAGD, [512, 32, 4, 2, 0], 10)                   # fused, fifth grid, NT
    with pytest.raises(smg.SMGError):
        smg.tune_set(smg.TUNE_DDAGD, [512, 32, 4, 2, 0], 11)               # no sixth grid
    smg.tune_clear()
    bad = str(tmp_path / "bad.txt")
    open(bad, "w").write("0 64 64 32 32 2 1\n0 64 64 32 32 x 1\n")
    with pytest.raises(smg.SMGError):
        smg.tune_load(bad)
    assert smg.tune_get(smg.TUNE_COARSE_BF16, [64, 64, 32, 32, 2]) == -1   # nothing added
