"""The random projection order (ProjectionSetRandom::reorder,
ettention/libReconstruction/src/model/projectionset/ProjectionSetRandom.cpp:17-38)
pinned part by part without Boost (CPU only):

* the engine: the oracle's mt19937 against libstdc++'s std::mt19937 (the same
  standard engine as boost::random::mt19937, default seed 5489), 1000 outputs;
* boost::uniform_int<>(0, r) where the mathematics fixes it: for r + 1 = 2^k
  the bucket is 2^(32-k) and nothing is rejected, so each draw is the top k bits
  of one engine output; for any r the buckets are floor(2^32 / (r + 1)) wide
  from the bottom and the partial top bucket is rejected (the only split into
  equal buckets that rejects least);
* the shuffle composition: libstdc++'s own std::random_shuffle, driven with the
  oracle's draws, gives exactly the oracle's order, which the library's order
  (projection_set.cpp) must equal as well.

What stays open: whether the reference build's Boost version (>= 1.55, not
pinned further) used this bucket rule; Boost has used it since 1.47.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")
SIZES = (1, 2, 3, 4, 5, 7, 8, 16, 61, 100, 118, 121, 255, 256, 1000)


@pytest.fixture(scope="module")
def libstdcxx(tmp_path_factory):
    O.lib()   # builds oracle/liboracle.so if needed
    exe = str(tmp_path_factory.mktemp("shuffle") / "std_random_shuffle")
    # C++14: std::random_shuffle with a generator is still in the standard there
    subprocess.run(["g++", "-O1", "-std=c++14", "-Wno-deprecated-declarations",
                    os.path.join(HERE, "native", "std_random_shuffle.cpp"), "-o", exe,
                    "-L" + ORACLE_DIR, "-loracle", "-Wl,-rpath," + ORACLE_DIR], check=True)
    out = subprocess.run([exe] + [str(n) for n in SIZES], check=True, capture_output=True, text=True).stdout
    lines = {}
    for line in out.splitlines():
        head, *vals = line.split()
        if head == "engine":
            lines["engine"] = [int(v) for v in vals]
        else:
            lines[int(vals[0])] = [int(v) for v in vals[1:]]
    return lines


def test_engine_matches_libstdcxx_mt19937(libstdcxx):
    m = O.Mt19937(5489)
    assert [m.next() for _ in range(1000)] == libstdcxx["engine"]


def test_uniform_int_power_of_two_ranges_take_the_top_bits(libstdcxx):
    eng = libstdcxx["engine"]
    m = O.Mt19937(5489)
    pos = 0
    for k in range(1, 17):
        for _ in range(20):
            got = m.uniform_int(0, (1 << k) - 1)
            assert got == eng[pos] >> (32 - k), (k, pos)
            pos += 1


def test_uniform_int_equal_buckets_from_the_bottom(libstdcxx):
    eng = libstdcxx["engine"]
    rng = np.random.default_rng(11)
    ranges = [1, 2, 3, 4, 5, 6, 7, 9, 60, 120, 1000, 65535, 65536, 3 * 2 ** 29] + \
        [int(r) for r in rng.integers(1, 2 ** 31, 40)]
    m = O.Mt19937(5489)
    pos = 0
    for r in ranges:
        bucket = 2 ** 32 // (r + 1)
        # the draw consumes engine outputs until one falls into a full bucket
        while True:
            x = eng[pos]
            pos += 1
            if x // bucket <= r:
                expect = x // bucket
                break
        assert m.uniform_int(0, r) == expect, r
    # a shifted range maps onto lo + the same draw
    a, b = O.Mt19937(5489), O.Mt19937(5489)
    for r in (1, 6, 120):
        assert a.uniform_int(10, 10 + r) == 10 + b.uniform_int(0, r)
    # an empty range consumes no engine output
    c, d = O.Mt19937(5489), O.Mt19937(5489)
    assert c.uniform_int(7, 7) == 7
    assert c.next() == d.next()


def test_oracle_order_is_libstdcxx_random_shuffle(libstdcxx):
    for n in SIZES:
        assert O.projection_order("random", n).tolist() == libstdcxx[n], n


def test_library_order_equals_the_pinned_order(libstdcxx):
    from ettention_amd import api as A
    for n in SIZES:
        assert A.projection_order("random", n).tolist() == libstdcxx[n], n
    # the skip list is ignored by the random scheme: the list restarts at 0..n-1
    assert A.projection_order("random", 121, skip=(0, 5, 9)).tolist() == libstdcxx[118]
