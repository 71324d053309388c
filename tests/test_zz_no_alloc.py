"""The allocation-free fill entry point (tmi_fill_residual_noalloc; run on an MI355X).

Called through the C ABI on two problems in a row, the second smaller, then the first again: F and
the sum of squares it writes to device memory equal tmi_fill_residual's bit for bit.  The first
problem fills in FFT form (the partial sums come from the fused last passes), the second in direct
1D form (the partial sums of one dot product), the third in the ATTInA22Block layout.  The scratch
array keeps the previous call's contents, so a stale partial sum would show.
"""
import ctypes as ct
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd import _lib  # noqa: E402
from tramonto_amd.problem import small_case  # noqa: E402


def state(spec, seed):
    o = Oracle.from_spec(spec)
    x = o.init_fields(spec.initial_x())
    x[spec.ncomp:] *= 1.0 + 0.01 * np.random.default_rng(seed).uniform(-1.0, 1.0, x[spec.ncomp:].shape)
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def both(dp, x):
    L = _lib.lib()
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    F1 = torch.zeros_like(x)
    s1 = ct.c_double()
    _lib.check(L.tmi_fill_residual(dp.handle, ct.c_void_p(x.data_ptr()), ct.c_void_p(F1.data_ptr()), ct.byref(s1),
                                   stream), "tmi_fill_residual")
    F2 = torch.full_like(x, np.nan)
    s2 = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    _lib.check(L.tmi_fill_residual_noalloc(dp.handle, ct.c_void_p(x.data_ptr()), ct.c_void_p(F2.data_ptr()),
                                           ct.c_void_p(s2.data_ptr()), stream), "tmi_fill_residual_noalloc")
    torch.cuda.synchronize()
    return F1.cpu().numpy(), s1.value, F2.cpu().numpy(), float(s2.item())


@pytest.mark.gpu
def test_noalloc_fill_equals_the_allocating_fill(monkeypatch):
    from tramonto_amd.device import DeviceProblem
    monkeypatch.setenv("TMI_FMT_FFT", "1")
    big = small_case("3d_periodic_small")
    dp_big = DeviceProblem(big)
    monkeypatch.delenv("TMI_FMT_FFT")
    assert dp_big.info.fmt_fft >= 1
    small = small_case("1d_hw")
    dp_small = DeviceProblem(small)
    a22 = dataclasses.replace(small_case("1d_reflect_wca"), att_in_a22=1)
    dp_a22 = DeviceProblem(a22)
    xb, xs = state(big, 1), state(small, 2)
    xa = state(dataclasses.replace(a22, att_in_a22=0), 3)[: a22.nunk].contiguous()
    for dp, x in ((dp_big, xb), (dp_small, xs), (dp_big, xb), (dp_a22, xa)):
        F1, s1, F2, s2 = both(dp, x)
        assert np.array_equal(F1, F2)
        assert s1 == s2, (s1, s2)
        assert abs(s1 - float((F1 ** 2).sum())) <= 1e-12 * s1
