"""The GA's random stream: its normals use the stream's own logarithm and
cosine (basic IEEE operations only, so host and device draw the same bits).
Pinned on the CPU: the oracle's restatement against the mathematics (Python's
math library within a few ulp, exact special values), and the library's
functions and normal draws against the oracle's, bit for bit."""
import math

import numpy as np

import oracle.oracle as orc
from gdock_amd import _lib


def _lib_math(x, which):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    _lib.lib().gdock_stream_math(_lib.ptr(x), len(x), which, _lib.ptr(out))
    return out


def _inputs():
    rng = np.random.default_rng(3)
    u = (rng.integers(1, 2 ** 53, 200000, dtype=np.int64).astype(np.float64)) * 2.0 ** -53  # the stream's grid
    wide = np.exp(rng.uniform(-700, 700, 20000))
    return u, wide


def test_ln_against_the_mathematics():
    u, wide = _inputs()
    for x in (u, wide):
        got = orc.stream_math(x, 0)
        ref = np.array([math.log(v) for v in x])
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
        assert err.max() <= 4.5e-16, err.max()
    exact = orc.stream_math(np.array([1.0, 2.0, 0.5, 4.0, 2.0 ** -53]), 0)
    assert exact[0] == 0.0
    assert exact[1] == math.log(2.0) and exact[2] == -math.log(2.0) and exact[3] == math.log(4.0)
    assert abs(exact[4] - (-53 * math.log(2.0))) <= 1e-14


def test_cos2pi_against_the_mathematics():
    u, _ = _inputs()
    got = orc.stream_math(u, 1)
    ref = np.cos(2.0 * np.pi * u)  # the argument's own rounding: |d cos| <= 2 pi u 1.1e-16
    assert np.abs(got - ref).max() <= 1.6e-15
    sp = orc.stream_math(np.array([0.0, 0.25, 0.5, 0.75, 0.125, 1.0 / 3.0]), 1)
    assert sp[0] == 1.0 and sp[1] == 0.0 and sp[2] == -1.0 and sp[3] == 0.0
    # within 3 ulp of the true values (cos(2 pi fl(1/3)) = -0.5 + 1.0e-16)
    assert abs(sp[4] - math.sqrt(0.5)) <= 3.4e-16 and abs(sp[5] - (-0.5 + 1.0e-16)) <= 3.4e-16


def test_library_stream_functions_equal_the_oracle_bitwise():
    u, wide = _inputs()
    for x, which in ((u, 0), (wide, 0), (u, 1)):
        assert np.array_equal(_lib_math(x, which).view(np.uint64), orc.stream_math(x, which).view(np.uint64))
    s = _lib_math(wide, 2)
    assert np.array_equal(s, np.sqrt(wide))  # correctly rounded


def test_normals_equal_the_oracle_bitwise():
    for seed in (0, 42, 7, 2 ** 63 + 5):
        n = 20000
        got = np.zeros(n)
        _lib.lib().gdock_rng_normals(seed, n, 0.0, 0.174, _lib.ptr(got))
        ref = orc.rng_normals(seed, n, 0.0, 0.174)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), seed
        z = got / 0.174
        assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.02
