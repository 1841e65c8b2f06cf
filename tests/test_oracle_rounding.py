"""The oracle's storage rounding (f32 -> bf16 round to nearest even, f32 ->
f16 round to nearest even) against library routines that fix it: torch's
CPU bfloat16 conversion and numpy's float16 conversion, on edge cases (ties
in both directions, denormals, overflow to infinity, signed zeros) and
random bit patterns.  NaN payloads are not compared: every rounding keeps
NaN a NaN."""
import numpy as np
import pytest

from oracle import oracle as O

torch = pytest.importorskip("torch")


def rounding_cases(n_random=20000, seed=3):
    edge = np.array([0x00000000, 0x80000000, 0x00000001, 0x007fffff, 0x00800000, 0x3f800000, 0x3f808000,
                     0x3f818000, 0x3f807fff, 0x3f808001, 0x3f817fff, 0x7f7f7fff, 0x7f7f8000, 0x7f7fffff,
                     0x7f800000, 0xff800000, 0x477fe000, 0x477ff000, 0x477fefff, 0x33800000, 0x33000000,
                     0x33000001, 0x387fc000, 0x80008000, 0x00018000, 0x00008000], dtype=np.uint32)
    rnd = np.random.default_rng(seed).integers(0, 2 ** 32, n_random, dtype=np.uint64).astype(np.uint32)
    bits = np.concatenate([edge, rnd])
    f = bits.view(np.float32)
    return f[~np.isnan(f)]


def test_bf16_rounding_matches_torch():
    f = rounding_cases()
    want = torch.from_numpy(f).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([O.lib().orc_f32_to_bf16_rne(float(x)) for x in f], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)


def test_f16_rounding_matches_numpy():
    f = rounding_cases()
    with np.errstate(over="ignore"):
        want = f.astype(np.float16).view(np.uint16)
    got = np.array([O.lib().orc_f32_to_f16_rte(float(x)) for x in f], dtype=np.uint16)
    np.testing.assert_array_equal(got, want)
