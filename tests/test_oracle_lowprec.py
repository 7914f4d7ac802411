"""Pins of the oracle's low-precision storage model (oracle/lowprec.py).

The FP8 E4M3 format is checked against values the OCP OFP8 specification
fixes (bias 7, largest finite 448, smallest normal 2^-6, smallest subnormal
2^-9, NaN S.1111.111) and the ties-to-even rule; bfloat16 against float32
bit patterns.  No GPU needed.
"""

import numpy as np
import pytest

from oracle import lowprec as lp


def test_e4m3_known_encodings():
    vals = np.array([1.0, 448.0, -448.0, 2.0 ** -6, 2.0 ** -9, 7 * 2.0 ** -9, 0.0, 1.75, 2.5,
                     -0.5], np.float32)
    codes = lp.e4m3_encode(lp.e4m3_quantize(vals))
    assert codes.tolist() == [0x38, 0x7E, 0xFE, 0x08, 0x01, 0x07, 0x00, 0x3E, 0x42, 0xB0]
    np.testing.assert_array_equal(lp.e4m3_decode(codes), vals)
    assert np.isnan(lp.e4m3_decode(np.array([0x7F, 0xFF], np.uint8))).all()


def test_e4m3_ties_to_even():
    x = np.array([1.0625, 1.1875, 2.0 ** -10, 3 * 2.0 ** -10, 240.0 + 8.0, 9.5 * 2.0 ** -9],
                 np.float32)
    # 1.0625 lies halfway between 1 and 1.125 -> 1 (even mantissa 0); 1.1875
    # halfway between 1.125 and 1.25 -> 1.25; half the smallest subnormal ->
    # 0; 1.5 subnormal quanta -> 2 quanta; 248 halfway between 240 and 256 ->
    # 256; 9.5 quanta of 2^-9 (normal range from 8 quanta: quantum 2^-9) -> 10
    want = np.array([1.0, 1.25, 0.0, 2 * 2.0 ** -9, 256.0, 10 * 2.0 ** -9], np.float32)
    np.testing.assert_array_equal(lp.e4m3_quantize(x), want)
    # just above the tie rounds up
    assert lp.e4m3_quantize(np.float32(np.nextafter(np.float32(1.0625), np.float32(2)))) == 1.125
    with pytest.raises(ValueError):
        lp.e4m3_quantize(np.float32(465.0))


def test_e4m3_round_trip_of_every_finite_code():
    codes = np.array([c for c in range(256) if c not in (0x7F, 0xFF)], np.uint8)
    vals = lp.e4m3_decode(codes)
    np.testing.assert_array_equal(lp.e4m3_quantize(vals), vals)
    back = lp.e4m3_encode(vals)
    # +0 and -0 both decode to zero; every other code comes back unchanged
    np.testing.assert_array_equal(back, codes)


def test_bf16_round_ties_to_even():
    x = np.array([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -3.0], np.float32)
    want = np.array([1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -3.0], np.float32)
    np.testing.assert_array_equal(lp.bf16_round(x), want)
    assert lp.bf16_bits(np.float32(1.0)) == 0x3F80
    # hi + lo recovers about 16 significant bits
    rng = np.random.default_rng(0)
    a = rng.standard_normal(10000).astype(np.float32)
    hi = lp.bf16_round(a)
    lo = lp.bf16_round(a - hi)
    assert np.max(np.abs((hi.astype(np.float64) + lo) - a) / np.abs(a)) < 2.0 ** -16


def test_scale_exponent_boundary():
    assert lp.scale_exponent(448.0) == 0
    assert lp.scale_exponent(np.nextafter(np.float32(448.0), np.float32(1e9))) == 1
    assert lp.scale_exponent(224.0) == -1
    assert lp.scale_exponent(0.0) == 0
    for m in [1.1, 3e-3, 7.0e4, 448.0 * 2.0 ** -20]:
        k = lp.scale_exponent(m)
        assert m * 2.0 ** -k <= 448.0 < m * 2.0 ** -(k - 1)


@pytest.mark.parametrize("storage", ["bf16x2", "bf16", "fp8x2", "fp8"])
def test_pack_model_layout_and_accuracy(storage):
    rng = np.random.default_rng(1)
    n = 70  # 3 x 3 tiles, the last ones zero-padded
    A = rng.standard_normal((n, n)).astype(np.float32)
    A[:32, :32] *= 1e-3  # a tile of small entries gets its own scale
    packed, scales, planes = lp.pack_model(A, storage)
    nt = 3
    per_tile = 2048 if storage.startswith("bf16") else 1024
    nplanes = 2 if storage.endswith("x2") else 1
    assert packed.shape == (nplanes * nt * nt * per_tile,)
    if storage.startswith("fp8"):
        assert scales.shape == (nplanes * nt * nt,)
        assert scales[0] < scales[1]  # the small tile's scale
    rep = lp.represented(A, storage)
    rel = np.abs(rep - A) / np.abs(A)
    bound = {"bf16x2": 2.0 ** -16, "bf16": 2.0 ** -8, "fp8x2": 2.0 ** -7, "fp8": 2.0 ** -3}[storage]
    assert np.max(rel) <= bound
    # the first tile's lane 33 (row 1, lane half 1), slot 9 is column 16 + 8 + 1
    if storage == "fp8":
        v = lp.e4m3_decode(packed[33 * 16 + 9:33 * 16 + 10])[0] * scales[0]
        assert v == planes[0][1, 25]
    if storage == "bf16":
        word = packed[2 * (512 + 33 * 8 + 1):2 * (512 + 33 * 8 + 1) + 2].view(np.uint16)[0]
        assert word == lp.bf16_bits(planes[0][1, 25])
