"""CPU check of the arithmetic behind NW_PREC_BF16X3: an FP32 value split
into three round-to-nearest-even BF16 parts keeps 24 significant bits, and
the six cross products h_a h_b + h_a m_b + m_a h_b + h_a l_b + l_a h_b + m_a m_b
(the terms the BF16 MFMA transform accumulates) reproduce the FP64 product of
the FP32 inputs to about 2^-24 relative, the level of one FP32 rounding."""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest, ties to even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def split3(v):
    v = np.asarray(v, dtype=np.float32)
    h = bf16(v)
    r = (v - h).astype(np.float32)
    m = bf16(r)
    lo = bf16((r - m).astype(np.float32))
    return h, m, lo


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    assert bf16(np.float32(1.0 + 2.0 ** -8)) == np.float32(1.0)
    assert bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == np.float32(1.0 + 2.0 ** -6)
    assert bf16(np.float32(-2.5)) == np.float32(-2.5)


def test_split_reconstructs_fp32():
    rng = np.random.default_rng(0)
    v = (rng.normal(size=100000) * 10.0 ** rng.uniform(-3, 3, size=100000)).astype(np.float32)
    h, m, lo = split3(v)
    rec = h.astype(np.float64) + m.astype(np.float64) + lo.astype(np.float64)
    rel = np.abs(rec - v.astype(np.float64)) / np.abs(v.astype(np.float64))
    assert rel.max() <= 2.0 ** -24


def test_six_term_product_error():
    rng = np.random.default_rng(1)
    a = rng.normal(size=200000).astype(np.float32)
    b = rng.normal(size=200000).astype(np.float32)
    ha, ma, la = (x.astype(np.float64) for x in split3(a))
    hb, mb, lb = (x.astype(np.float64) for x in split3(b))
    six = ha * hb + ha * mb + ma * hb + ha * lb + la * hb + ma * mb
    exact = a.astype(np.float64) * b.astype(np.float64)
    rel = np.abs(six - exact) / np.abs(exact)
    assert rel.max() < 4 * 2.0 ** -24
    # a two-way split (three products) is far coarser: the claim the six terms rest on
    h2a, m2a = bf16(a).astype(np.float64), bf16((a - bf16(a)).astype(np.float32)).astype(np.float64)
    h2b, m2b = bf16(b).astype(np.float64), bf16((b - bf16(b)).astype(np.float32)).astype(np.float64)
    two = h2a * h2b + h2a * m2b + m2a * h2b
    assert (np.abs(two - exact) / np.abs(exact)).max() > 2.0 ** -18
