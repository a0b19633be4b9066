"""pbf_step_noalloc (the entry point that allocates nothing: every temporary
in the library's static device scratch) against the handle path
pbf_create + pbf_step + pbf_dump on the same inputs: bit-identical results,
over calls in a row on one GPU whose later inputs are smaller (the shared
scratch keeps the last call's contents)."""
import ctypes

import numpy as np
import pytest

import pbf
import scenarios

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    pbf.load_library()  # loads the HIP runtime the library uses
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                              ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    for f in ("hipMalloc", "hipMemcpy", "hipFree", "hipDeviceSynchronize"):
        getattr(hip, f).restype = ctypes.c_int
    return hip


def _noalloc(p, parts, nsteps):
    """The state after nsteps steps through pbf_step_noalloc on device
    buffers the test allocates itself."""
    hip, lib = _hip(), pbf.load_library()
    n = parts.shape[0]
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, :3] = parts[:, 1:4]
    vel[:, :3] = parts[:, 4:7]
    vel[:, 3] = parts[:, 0]
    dp, dv = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dp), pos.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dv), vel.nbytes) == 0
    try:
        assert hip.hipMemcpy(dp, pos.ctypes.data, pos.nbytes, H2D) == 0
        assert hip.hipMemcpy(dv, vel.ctypes.data, vel.nbytes, H2D) == 0
        rc = lib.pbf_step_noalloc(ctypes.byref(p), dp, dv, n, nsteps, None)
        assert rc == 0, lib.pbf_last_error().decode()
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(pos.ctypes.data, dp, pos.nbytes, D2H) == 0
        assert hip.hipMemcpy(vel.ctypes.data, dv, vel.nbytes, D2H) == 0
    finally:
        hip.hipFree(dp)
        hip.hipFree(dv)
    return pos, vel


def _handle(p, parts, nsteps):
    with pbf.Simulation(p, parts) as s:
        s.step(nsteps)
        return s.dump()


def _case(dims, physics, seed, mixed=False):
    sc = scenarios.small_dam(dims=dims, box=(2.0, 2.0, 2.0), physics=physics,
                             seed=seed)
    p = sc.params[0]
    if mixed:
        p.precision, p.fp32_iterations = pbf.PRECISION_MIXED, 1
    return p, sc.particles[0]


def test_noalloc_matches_handle_over_calls_in_a_row():
    # each later input smaller than the one before: a call must not depend
    # on what an earlier call left in the scratch
    cases = [((16, 16, 16), pbf.PHYSICS_PAPER, 41, False, 5),
             ((10, 10, 10), pbf.PHYSICS_PAPER, 42, False, 5),
             ((9, 9, 9), pbf.PHYSICS_REF, 43, False, 3),
             ((8, 8, 8), pbf.PHYSICS_PAPER, 44, True, 4)]
    for dims, physics, seed, mixed, nsteps in cases:
        p, parts = _case(dims, physics, seed, mixed)
        gp, gv = _noalloc(p, parts, nsteps)
        hp, hv = _handle(p, parts, nsteps)
        assert np.isfinite(gp).all() and np.abs(gp[:, :3] - parts[:, 1:4]).max() > 0
        assert np.array_equal(gp, hp), (dims, physics, mixed)
        assert np.array_equal(gv, hv), (dims, physics, mixed)


def test_noalloc_refuses_what_does_not_fit():
    lib = pbf.load_library()
    p, parts = _case((8, 8, 8), pbf.PHYSICS_PAPER, 45)
    dummy = (ctypes.c_float * 8)()
    rc = lib.pbf_step_noalloc(ctypes.byref(p), dummy, dummy, 1_000_000, 1, None)
    assert rc == pbf.ERR_INVALID
    assert "scratch" in lib.pbf_last_error().decode()
