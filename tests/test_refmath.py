"""CPU pins of the FMA sequences behind the REF pair terms' division and
square root (position-based-fluids_amd/csrc/pbf_refmath.hpp).

The REF passes reproduce the reference kernels' IEEE results bit for bit
(`src/kernels/compute_scaling.cl:62-67`, `src/kernels/compute_delta.cl:59-60`:
a square root and three divisions per pair;
`src/kernels/apply_vorticity_and_viscosity.cl:61`: 1/rho_j).
The device code forms them with Newton-Raphson steps from the hardware's
reciprocal and square-root estimates (accurate to 1 ulp).  These tests run
the same steps on the host with C `fmaf` (tests/native/refmath_cpu.c),
started from EVERY float within 1 ulp of the exact value, and compare with
IEEE division and sqrtf: exhaustively over a binade pair for the square
root (the result depends on the significand and the exponent's parity
only), over a binade for the reciprocal, and on random and
round-to-midpoint cases for the division.  The GPU tests
(test_gpu_refmath.py) compare the device code with the compiler's general
sequences on the hardware.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
P = ctypes.POINTER(ctypes.c_float)


@pytest.fixture(scope="module")
def lib():
    path = os.path.join(NATIVE, "librefmath_cpu.so")
    if not os.path.exists(path):
        subprocess.run(["make", "-C", NATIVE, "librefmath_cpu.so"], check=True)
    lib = ctypes.CDLL(path)
    lib.check_div.argtypes = [P, P, ctypes.c_long,
                              ctypes.POINTER(ctypes.c_long)]
    lib.check_div.restype = ctypes.c_long
    for f in (lib.check_recip, lib.check_sqrt):
        f.argtypes = [P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]
        f.restype = ctypes.c_long
    return lib


def _run(fn, *arrays):
    first = ctypes.c_long()
    arrays = [np.ascontiguousarray(a, dtype=np.float32) for a in arrays]
    bad = fn(*[a.ctypes.data_as(P) for a in arrays], len(arrays[0]),
             ctypes.byref(first))
    return bad, first.value


def _binade(lo_bits, hi_bits):
    return np.arange(lo_bits, hi_bits, dtype=np.uint32).view(np.float32)


def test_sqrt_steps_exhaustive(lib):
    # every x in [1, 4): all significands at both exponent parities
    assert _run(lib.check_sqrt, _binade(0x3F800000, 0x40800000)) == (0, -1)
    # the admitted range's ends and random values across it
    rng = np.random.default_rng(11)
    x = np.concatenate([
        _binade(0x0F800000, 0x0F800000 + (1 << 20)),       # from 2^-96 up
        _binade(0x7F7FFFFF - (1 << 20), 0x7F800000),       # to FLT_MAX
        np.exp2(rng.uniform(-96, 127.99, 1 << 20)).astype(np.float32),
    ])
    assert _run(lib.check_sqrt, x) == (0, -1)


def test_recip_steps_exhaustive(lib):
    """From an estimate within 1 ulp the steps give 1/b correctly rounded
    for every b in [1, 2) but one: the all-ones significand 2 - 2^-23,
    whose reciprocal lies 2^-49 above a rounding midpoint.  There only the
    correctly rounded estimate itself converges (the two neighbours end on
    the even side, 0.5): the known hard case of Newton-Raphson reciprocals.
    The GPU test compares the device steps with the compiler's 1/b on
    hardware for this significand in several binades."""
    b = _binade(0x3F800000, 0x40000000)
    bad, first = _run(lib.check_recip, b)
    assert (bad, first) == (2, (1 << 23) - 1)
    rng = np.random.default_rng(12)
    r = np.exp2(rng.uniform(-90, 90, 1 << 21)).astype(np.float32)
    r = r[(r.view(np.uint32) & 0x7FFFFF) != 0x7FFFFF]
    assert _run(lib.check_recip, r) == (0, -1)


def test_div_steps_random_and_midpoints(lib):
    rng = np.random.default_rng(13)
    n = 1 << 21
    # divisors over the admitted range, numerators |a| <= b of either sign,
    # down to the smallest admitted magnitude 2^-103
    b = np.exp2(rng.uniform(-48, 20, n)).astype(np.float32)
    a = (b * rng.uniform(-1, 1, n)).astype(np.float32)
    assert _run(lib.check_div, a, b) == (0, -1)
    tiny = (np.where(rng.random(n) < 0.5, -1.0, 1.0)
            * np.exp2(rng.uniform(-103, -60, n))).astype(np.float32)
    assert _run(lib.check_div, tiny, b) == (0, -1)
    zeros = np.where(rng.random(n) < 0.5, -0.0, 0.0).astype(np.float32)
    assert _run(lib.check_div, zeros, b) == (0, -1)
    # quotients next to a rounding midpoint: a = RN(b m), m halfway between
    # two floats of [1, 2); also the numerators one ulp either side
    bs = _binade(0x3F800000, 0x40000000)[rng.integers(0, 1 << 23, n)]
    q = rng.integers(0x3F800000, 0x40000000, n, dtype=np.uint32).view(
        np.float32)
    m = q.astype(np.float64) + np.spacing(q).astype(np.float64) / 2
    am = (bs.astype(np.float64) * m).astype(np.float32)
    for d in (0, 1, -1):
        ad = (am.view(np.uint32) + np.uint32(d & 0xFFFFFFFF)).view(np.float32)
        assert _run(lib.check_div, ad, bs) == (0, -1)
