"""Pins of the oracle's scalar numerics against independent references:
libm exp, published SplitMix64 outputs, numpy's float16 and torch's bfloat16
round-to-nearest-even conversions."""
import ctypes as C
import math
import struct

import numpy as np
import pytest

import oracle as O
from lattibol_amd import scenes


def ulp_diff(a, b):
    ia = struct.unpack("<q", struct.pack("<d", a))[0]
    ib = struct.unpack("<q", struct.pack("<d", b))[0]
    return abs(ia - ib)


def test_exp_matches_libm_within_2ulp():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 10000), rng.uniform(-2, 2, 10000),
                         [-300.0 / 1e9, 0.0, 1.0, -1.0, math.log(2.0), -200.0 / 524.39, -200.0 / 85.7]])
    worst = max(ulp_diff(O.lib.orc_exp(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 2, worst
    assert O.lib.orc_exp(0.0) == 1.0
    assert abs(O.lib.orc_exp(math.log(2.0)) - 2.0) <= 4.5e-16
    assert O.lib.orc_exp(-750.0) == 0.0


def test_splitmix64_published_outputs():
    # SplitMix64 with state 0 (Vigna's reference splitmix64.c): first outputs
    st = C.c_uint64(0)
    got = [O.lib.orc_splitmix64_next(C.byref(st)) for _ in range(3)]
    assert got[0] == 0xE220A8397B1DCDAF
    assert got[1] == 0x6E789E6AA1B965F4
    assert got[2] == 0x06C45D188009454F


def test_product_generator_equals_oracle_generator():
    seed = 12345
    st = C.c_uint64(seed)
    seq = [O.lib.orc_splitmix64_next(C.byref(st)) for _ in range(64)]
    assert [int(v) for v in scenes.splitmix64(seed, 64)] == seq
    assert [int(v) for v in scenes.splitmix64(seed, 10, start=20)] == seq[20:30]
    out = np.empty(1000, np.float32)
    O.lib.orc_uniform_noise(7, 1000, 0.7, 0.01, O.P(out))
    np.testing.assert_array_equal(out, scenes.uniform_noise(7, 1000, 0.7, 0.01))
    assert out.min() >= np.float32(0.69) and out.max() <= np.float32(0.71)


def _edge_floats():
    rng = np.random.default_rng(1)
    vals = [0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 70000.0, 6.1e-5, 6.0e-8, 5.96e-8, 2.98e-8,
            3.0e-8, 1e-10, np.inf, -np.inf]
    # exact ties between half values (round to even)
    halves = np.arange(1, 4096, dtype=np.uint16).view(np.float16).astype(np.float32)
    ties = (halves[:-1].astype(np.float64) + halves[1:].astype(np.float64)) / 2
    return np.concatenate([np.array(vals, np.float32), rng.standard_normal(20000).astype(np.float32) * 0.05,
                           rng.uniform(-1, 1, 20000).astype(np.float32) * 1e-5, ties.astype(np.float32),
                           -ties.astype(np.float32)])


def test_f16_conversion_is_numpy_rne():
    x = _edge_floats()
    out = np.empty(x.size, np.uint16)
    O.lib.orc_f32_to_f16_array(O.P(x), O.P(out), x.size)
    np.testing.assert_array_equal(out, x.astype(np.float16).view(np.uint16))
    back = np.array([O.lib.orc_f16_to_f32(int(h)) for h in out[:2000]], np.float32)
    np.testing.assert_array_equal(back, out[:2000].view(np.float16).astype(np.float32))


def test_bf16_conversion_is_torch_rne():
    torch = pytest.importorskip("torch")
    x = _edge_floats()
    out = np.empty(x.size, np.uint16)
    O.lib.orc_f32_to_bf16_array(O.P(x), O.P(out), x.size)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    np.testing.assert_array_equal(out, ref)


def test_residual_norm_ratio():
    rng = np.random.default_rng(3)
    n = 500
    a = rng.standard_normal((3, n)).astype(np.float32)
    b = a + rng.standard_normal((3, n)).astype(np.float32) * 1e-3
    solid = (rng.uniform(size=n) < 0.1).astype(np.uint8)
    got = O.lib.orc_residual(3, n, O.P(a), O.P(b), O.P(solid))
    m = solid == 0
    ref = math.sqrt(math.fsum(((a[:, m].astype(np.float64) - b[:, m]) ** 2).ravel())
                    / math.fsum((a[:, m].astype(np.float64) ** 2).ravel()))
    assert abs(got - ref) <= 1e-12 * ref
