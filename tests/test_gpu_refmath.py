"""GPU check of the REF pair terms' division and square root
(position-based-fluids_amd/csrc/pbf_refmath.hpp) against the compiler's
general IEEE sequences on the hardware: every result must be the same bits
as '/', sqrtf and 1/x give (tests/native/refmath_gpu.hip), over whole
binades and the edges of the ranges where the shorter sequences apply, and
the range tests must classify their boundary values as stated.  The REF
passes use these helpers in every pair (`src/kernels/compute_scaling.cl:62-67`,
`src/kernels/compute_delta.cl:59-60`,
`src/kernels/apply_vorticity_and_viscosity.cl:61`).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NATIVE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native")
DIV, RECIP, SQRT, UNIT, REF_SQRT, REF_RECIP, G_TINY, G_SQRT, G_RECIP = range(9)


@pytest.fixture(scope="module")
def check():
    path = os.path.join(NATIVE, "librefmath_gpu.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (make -C tests/native)")
    lib = ctypes.CDLL(path)
    lib.refmath_check.argtypes = [ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_float),
                                  ctypes.c_long,
                                  ctypes.POINTER(ctypes.c_longlong)]
    lib.refmath_check.restype = ctypes.c_longlong

    def run(which, data):
        data = np.ascontiguousarray(data, dtype=np.float32)
        per = 3 if which == UNIT else 2 if which in (DIV, G_TINY, G_SQRT,
                                                     G_RECIP) else 1
        n = data.size // per
        print(f"refmath check {which}: {n} inputs", flush=True)
        first = ctypes.c_longlong()
        bad = lib.refmath_check(
            which, data.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), n,
            ctypes.byref(first))
        assert bad >= 0, "HIP error in the check"
        if bad:
            row = data.reshape(n, per)[first.value] if per > 1 else \
                data[first.value]
            return bad, row
        return 0, None
    return run


def _binade(lo_bits, hi_bits):
    return np.arange(lo_bits, hi_bits, dtype=np.uint32).view(np.float32)


def _f(bits):
    return np.asarray(bits, dtype=np.uint32).view(np.float32)


def test_sqrt_matches_sqrtf(check):
    rng = np.random.default_rng(21)
    x = np.concatenate([
        _binade(0x3F800000, 0x40800000),                    # [1, 4)
        _binade(0x0F800000, 0x10800000),                    # [2^-96, 2^-94)
        _binade(0x7E800000, 0x7F800000),                    # [2^126, inf)
        np.exp2(rng.uniform(-96, 127.99, 1 << 22)).astype(np.float32),
    ])
    assert check(SQRT, x) == (0, None)


def test_recip_matches_division(check):
    b = np.concatenate([
        _binade(0x3F800000, 0x40000000),                    # [1, 2)
        _binade(0x45800000, 0x46000000),                    # [4096, 8192)
        _binade(0x12800000, 0x13000000),                    # [2^-90, 2^-89)
        # every binade's all-ones significand
        _f(np.arange(37, 218, dtype=np.uint32) << 23 | 0x7FFFFF),
        _f([0x6C800000]),                                    # 2^90
    ])
    assert check(RECIP, b) == (0, None)


def test_div_matches_division(check):
    rng = np.random.default_rng(22)
    bs = _binade(0x3F800000, 0x40000000)                    # every b in [1, 2)
    n = bs.size
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    a = (sign * rng.uniform(1, 2, n) * np.exp2(rng.integers(-40, 1, n))
         ).astype(np.float32)
    assert check(DIV, np.stack([a, bs], 1)) == (0, None)
    # quotients next to rounding midpoints
    q = rng.integers(0x3F800000, 0x40000000, n, dtype=np.uint32).view(
        np.float32)
    m = q.astype(np.float64) + np.spacing(q).astype(np.float64) / 2
    am = (bs.astype(np.float64) * m).astype(np.float32)
    assert check(DIV, np.stack([am, bs], 1)) == (0, None)
    # divisors across the admitted range; numerators down to 2^-103, zeros
    b = np.exp2(rng.uniform(-48, 20, 1 << 22)).astype(np.float32)
    k = b.size
    a = (b * rng.uniform(-1, 1, k)).astype(np.float32)
    tiny = (np.where(rng.random(k) < 0.5, -1.0, 1.0)
            * np.exp2(rng.uniform(-103, -60, k))).astype(np.float32)
    tiny[: k // 8] = np.where(rng.random(k // 8) < 0.5, -0.0, 0.0)
    assert check(DIV, np.stack([a, b], 1)) == (0, None)
    assert check(DIV, np.stack([tiny, b], 1)) == (0, None)


def test_pair_unit_vector_matches_ieee(check):
    """ref_unit (the guarded |r| and r/|r| of a pair) against sqrtf and
    '/', on vectors of every scale, with zero, negative-zero and tiny
    components (which take the general sequences)."""
    rng = np.random.default_rng(23)
    n = 1 << 22
    r = (np.where(rng.random((n, 3)) < 0.5, -1.0, 1.0)
         * np.exp2(rng.uniform(-60, 19, (n, 1)))
         * rng.uniform(0, 1, (n, 3))).astype(np.float32)
    z = rng.random((n, 3))
    r[z < 0.05] = 0.0
    r[(z >= 0.05) & (z < 0.07)] = -0.0
    tiny = (z >= 0.07) & (z < 0.09)
    r[tiny] = (np.exp2(rng.uniform(-149, -95, tiny.sum()))).astype(np.float32)
    r[np.all(r == 0, axis=1), 0] = 0.5
    assert check(UNIT, r) == (0, None)


def test_guarded_sqrt_and_recip_any_input(check):
    rng = np.random.default_rng(24)
    special = _f([0, 0x80000000, 1, 0x007FFFFF, 0x00800000, 0x0F7FFFFF,
                  0x0F800000, 0x0F800001, 0x12800000, 0x127FFFFF,
                  0x6C800000, 0x6C800001, 0x7F7FFFFF, 0x7F800000,
                  0xFF800000, 0x7FC00000, 0xBF800000, 0x3FFFFFFF])
    bits = rng.integers(0, 1 << 32, 1 << 22, dtype=np.uint64).astype(
        np.uint32)
    x = np.concatenate([special, _f(bits)])
    assert check(REF_SQRT, x) == (0, None)
    assert check(REF_RECIP, x) == (0, None)


def test_range_tests_classify_their_edges(check):
    def labelled(values, pred):
        v = np.asarray(values, dtype=np.float32)
        return np.stack([v, pred(v.astype(np.float64)).astype(np.float32)], 1)

    edge = _f([0, 0x80000000, 1, 0x80000001, 0x007FFFFF, 0x00800000,
               0x0BFFFFFF, 0x0C000000, 0x8BFFFFFF, 0x8C000000, 0x0F7FFFFF,
               0x0F800000, 0x12800000, 0x127FFFFF, 0x6C800000, 0x6C800001,
               0x7F7FFFFF, 0x7F800000, 0xFF800000, 0x7FC00000, 0xBF800000,
               0x3F800000])
    rng = np.random.default_rng(25)
    rand = _f(rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(
        np.uint32))
    v = np.concatenate([edge, rand])
    with np.errstate(invalid="ignore"):
        assert check(G_TINY, labelled(
            v, lambda x: (x == 0) | (np.abs(x) >= 2.0 ** -103)
            | np.isnan(x))) == (0, None)
        assert check(G_SQRT, labelled(
            v, lambda x: (x >= 2.0 ** -96) & np.isfinite(x))) == (0, None)
        assert check(G_RECIP, labelled(
            v, lambda x: (x >= 2.0 ** -90) & (x <= 2.0 ** 90))) == (0, None)
