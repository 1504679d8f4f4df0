"""smg_dot_dev_noalloc, the allocation-free variant of smg_dot_dev: its partial rows,
completion counter and result live in the library's file-scope device scratch array.
Two calls in a row, the second on a smaller input, give bit-for-bit smg_dot_dev's
result (the same kernel, grid and summation order), and both agree with numpy."""
import ctypes

import numpy as np
import pytest

from helpers import banner

pytestmark = pytest.mark.gpu


def _dot(smg, fn, a, b, prec, stream):
    out = np.zeros(2)
    smg._chk(getattr(smg.lib(), fn)(a.data_ptr(), b.data_ptr(), a.numel(), prec,
                                    out.ctypes.data_as(ctypes.c_void_p), stream))
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_dot_equals_dot_dev(smg, prec):
    import torch
    banner(f"allocation-free dot {prec}")
    dt = torch.complex128 if prec == "f64" else torch.complex64
    p = smg.F64 if prec == "f64" else smg.F32
    stream = torch.cuda.current_stream().cuda_stream
    g = torch.Generator(device="cpu").manual_seed(7)
    for n in (1 << 20, 12345, 1):    # the second and third calls are smaller
        a = torch.randn(n, dtype=dt, generator=g).cuda()
        b = torch.randn(n, dtype=dt, generator=g).cuda()
        torch.cuda.synchronize()
        got = _dot(smg, "smg_dot_dev_noalloc", a, b, p, stream)
        want = _dot(smg, "smg_dot_dev", a, b, p, stream)
        assert np.array_equal(got, want), (n, got, want)
        ref = np.vdot(a.cpu().numpy().astype(np.complex128), b.cpu().numpy().astype(np.complex128))
        tol = 1e-12 if prec == "f64" else 1e-6
        assert abs(complex(got[0], got[1]) - ref) <= tol * max(1.0, np.sqrt(n)) * 4, (n, got, ref)


def test_noalloc_dot_refuses_bad_arguments(smg):
    banner("allocation-free dot argument checks")
    out = np.zeros(2)
    rc = smg.lib().smg_dot_dev_noalloc(None, None, 4, smg.F64, out.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == -1
    assert b"null argument" in smg.lib().smg_last_error()
