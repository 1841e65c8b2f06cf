"""The allocation-free residual norm (etn_sum_squared_difference_f64_noalloc)
against the allocating entry point it varies: the same kernels, so the same
FP64 result bit for bit, on two calls in a row (the second smaller, so stale
partial sums of the first would show) and on a stream of its own."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import ettention_amd as E  # noqa: E402
from ettention_amd import _lib as L  # noqa: E402



def _ssd(name, a, b, out, stream):
    lib = E.lib()
    st = getattr(lib, name)(C.c_void_p(a.data_ptr()), None if b is None else C.c_void_p(b.data_ptr()),
                            C.c_size_t(a.numel()), C.c_void_p(out.data_ptr()), C.c_void_p(stream))
    assert st == L.ETN_OK, lib.etn_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("on_side_stream", [False, True])
def test_noalloc_matches_allocating_entry_point(on_side_stream):
    rng = np.random.default_rng(7)
    side = torch.cuda.Stream() if on_side_stream else None
    stream = side.cuda_stream if side is not None else torch.cuda.current_stream().cuda_stream
    # first call large, second smaller (and without b): every partial of the
    # second call must be its own, not the first call's left in the scratch
    for count, with_b in ((3_000_001, True), (1_001, False), (77, True)):
        a_h = rng.normal(size=count).astype(np.float32)
        b_h = rng.normal(size=count).astype(np.float32) if with_b else None
        a = torch.from_numpy(a_h).cuda()
        b = torch.from_numpy(b_h).cuda() if with_b else None
        got = torch.full((1,), 0.25, dtype=torch.float64, device="cuda")   # the result accumulates into out
        ref = torch.full((1,), 0.25, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _ssd("etn_sum_squared_difference_f64_noalloc", a, b, got, stream)
        _ssd("etn_sum_squared_difference_f64", a, b, ref, stream)
        torch.cuda.synchronize()
        g, r = float(got.cpu()[0]), float(ref.cpu()[0])
        assert g == r, (count, g, r)
        d = a_h.astype(np.float64) - (b_h.astype(np.float64) if with_b else 0.0)
        expect = 0.25 + float((d * d).sum())
        assert abs(g - expect) <= 1e-12 * expect, (count, g, expect)


@pytest.mark.gpu
def test_noalloc_python_wrapper():
    from ettention_amd import api as A
    a = torch.arange(1000, dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    A.sum_squared_difference_noalloc(a, None, out)
    torch.cuda.synchronize()
    assert float(out.cpu()[0]) == float(sum(float(i) * i for i in range(1000)))


def test_noalloc_rejects_counts_above_the_maximum():
    """Host-side check, before any device call (runs without a GPU): the
    pointers are never dereferenced."""
    lib = E.lib()
    st = lib.etn_sum_squared_difference_f64_noalloc(C.c_void_p(16), None, C.c_size_t((1 << 40) + 1),
                                                     C.c_void_p(16), None)
    assert st == L.ETN_ERROR_INVALID_ARGUMENT
    assert b"maximum" in lib.etn_last_error()
