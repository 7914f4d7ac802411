"""Low-precision stored-matrix formats: NumPy model of the packed storage.

Test infrastructure (the oracle): restates, from the published format
definitions, what slq_pack_matrix stores for a stored matrix, so that the
GPU's packed bytes and its block matvec on them can be checked exactly.
Nothing in the product imports this file.

* BF16: IEEE binary32 truncated to its upper 16 bits after rounding to
  nearest, ties to even (bfloat16).
* FP8 E4M3 as in the OCP 8-bit Floating Point Specification (OFP8, v1.0):
  1 sign, 4 exponent (bias 7), 3 mantissa bits; subnormals for exponent
  field 0 (value m 2^-9); largest finite value 448 = 1.75 2^8; no
  infinities; S.1111.111 is NaN.  Conversion from FP32 rounds to nearest,
  ties to even.
* Tile scales: per 32 x 32 tile and plane the power of two 2^k with the
  smallest k such that max|tile| 2^-k <= 448; the stored bytes are the E4M3
  values of A 2^-k; the second plane (fp8x2) quantizes the exact residual
  A - 2^k q_hi of the tile the same way.
* Layout (include/slq.h, slq_storage): the matrix zero-padded to
  npad = 32 ceil(n / 32), tiles (rb, ct) at rb * (npad / 32) + ct, planes
  one after the other; inside a tile lane l = r + 32 h holds row r and the
  16 columns frag_k(s, h) = 8 h + s (s < 8), 16 + 8 h + (s - 8) (s >= 8);
  BF16 planes as [half][lane][8], FP8 planes as [lane][16].

Pinned by tests/test_oracle_lowprec.py: known E4M3 encodings (1, 448, the
smallest normal and subnormal values), ties-to-even cases of both formats,
the round trip of all 254 finite E4M3 codes, bfloat16 against the bit
pattern of float32 values that need no rounding, and the scale rule at its
boundary (448 and the next float above it).
"""

import numpy as np

TILE = 32
E4M3_MAX = 448.0


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    b = x.view(np.uint32).astype(np.uint64)
    rounded = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32).reshape(x.shape)


def bf16_bits(x):
    """The 16-bit patterns of bfloat16 values given as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    return (x.view(np.uint32) >> 16).astype(np.uint16)


def e4m3_decode(codes):
    """uint8 E4M3 codes -> float32 values (NaN for S.1111.111)."""
    c = np.asarray(codes, dtype=np.uint8).astype(np.int64)
    sign = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 0xF
    m = c & 0x7
    val = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * np.exp2(e - 7.0))
    val = np.where((e == 0xF) & (m == 0x7), np.nan, val)
    return (sign * val).astype(np.float32)


def e4m3_quantize(x):
    """float32 values with |x| <= 448 -> nearest E4M3 value (ties to even)."""
    x = np.asarray(x, dtype=np.float64)
    a = np.abs(x)
    # quantum: 2^(E - 3) in the normal range (E = floor(log2 a) >= -6), 2^-9 below
    _, e = np.frexp(np.where(a > 0, a, 1.0))  # a = f 2^e, f in [0.5, 1): E = e - 1
    quantum = np.exp2(np.maximum(e - 1, -6) - 3.0)
    q = np.round(a / quantum) * quantum  # numpy rounds halves to even
    if np.any(q > E4M3_MAX):
        raise ValueError("value beyond the E4M3 range")
    return (np.sign(x) * q).astype(np.float32)


def e4m3_encode(v):
    """E4M3 values (as produced by e4m3_quantize) -> uint8 codes."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    sign = np.where(np.signbit(v), 0x80, 0)
    _, e = np.frexp(np.where(a > 0, a, 1.0))
    E = e - 1
    normal = a >= 2.0 ** -6
    mant_n = np.round((a / np.exp2(E) - 1.0) * 8.0).astype(np.int64)
    code_n = ((E + 7) << 3) | mant_n
    code_s = np.round(a / 2.0 ** -9).astype(np.int64)
    code = np.where(normal, code_n, code_s)
    return (sign | code).astype(np.uint8)


def scale_exponent(m):
    """Smallest k with m 2^-k <= 448 (0 for m = 0), as slq_pack_matrix."""
    m = np.float32(m)
    if not m > 0:
        return 0
    f, e = np.frexp(m)
    k = int(e) - 9 if f <= 0.875 else int(e) - 8
    return max(k, -120)


def frag_columns():
    """(64, 16) column offsets within a tile: lane l, slot s."""
    lane = np.arange(64)[:, None]
    s = np.arange(16)[None, :]
    h = lane >> 5
    return np.where(s < 8, 8 * h + s, 16 + 8 * h + (s - 8))


def _tiles(A):
    """(nt, nt, 64, 16) fragment view of the zero-padded matrix."""
    A = np.asarray(A, dtype=np.float32)
    n = A.shape[0]
    nt = -(-n // TILE)
    Ap = np.zeros((nt * TILE, nt * TILE), np.float32)
    Ap[:n, :n] = A
    T = Ap.reshape(nt, TILE, nt, TILE).transpose(0, 2, 1, 3)  # (rb, ct, r, c)
    lane = np.arange(64)
    rows = (lane & 31)[:, None]
    cols = frag_columns()
    return T[:, :, rows, cols], nt  # (rb, ct, 64, 16)


def _untile(F, n):
    """Inverse of _tiles: (nt, nt, 64, 16) -> (n, n)."""
    nt = F.shape[0]
    T = np.zeros((nt, nt, TILE, TILE), F.dtype)
    lane = np.arange(64)
    rows = (lane & 31)[:, None]
    cols = frag_columns()
    T[:, :, rows, cols] = F
    return T.transpose(0, 2, 1, 3).reshape(nt * TILE, nt * TILE)[:n, :n]


def _fp8_plane(F):
    """Quantize each tile of F with its own scale: (values, codes, scales)."""
    nt = F.shape[0]
    codes = np.empty(F.shape, np.uint8)
    vals = np.empty(F.shape, np.float32)
    scales = np.empty((nt, nt), np.float32)
    for rb in range(nt):
        for ct in range(nt):
            t = F[rb, ct]
            k = scale_exponent(np.max(np.abs(t)))
            q = e4m3_quantize(np.ldexp(t, -k).astype(np.float32))
            codes[rb, ct] = e4m3_encode(q)
            vals[rb, ct] = np.ldexp(q, k)
            scales[rb, ct] = np.float32(2.0 ** k)
    return vals, codes, scales


def pack_model(A, storage):
    """Expected packed bytes, scales and the represented matrix of A.

    Returns ``(packed uint8 (bytes,), scales float32 (count,) or None,
    planes [float32 (n, n) per plane])``; the represented matrix is the sum of
    the planes (hi + lo, or hi alone)."""
    F, nt = _tiles(A)
    n = np.asarray(A).shape[0]
    if storage in ("bf16x2", "bf16"):
        hi = bf16_round(F)
        planes = [hi]
        if storage == "bf16x2":
            planes.append(bf16_round((F - hi).astype(np.float32)))
        out = []
        for p in planes:
            # per tile [half][lane][8]: slots 0-7, then slots 8-15
            bits = bf16_bits(p)  # (rb, ct, 64, 16)
            halves = np.stack([bits[..., :8], bits[..., 8:]], axis=2)  # (rb, ct, 2, 64, 8)
            out.append(halves.reshape(-1).view(np.uint8))
        return (np.concatenate(out), None, [_untile(p, n) for p in planes])
    if storage in ("fp8x2", "fp8"):
        v_hi, c_hi, s_hi = _fp8_plane(F)
        vals, codes, scales = [v_hi], [c_hi], [s_hi]
        if storage == "fp8x2":
            v_lo, c_lo, s_lo = _fp8_plane((F - v_hi).astype(np.float32))
            vals.append(v_lo)
            codes.append(c_lo)
            scales.append(s_lo)
        packed = np.concatenate([c.reshape(-1) for c in codes])
        return (packed, np.concatenate([s.reshape(-1) for s in scales]),
                [_untile(v, n) for v in vals])
    raise ValueError(f"unknown storage {storage!r}")


def represented(A, storage):
    """The matrix the packed storage represents (float64)."""
    _, _, planes = pack_model(A, storage)
    return sum(p.astype(np.float64) for p in planes)
