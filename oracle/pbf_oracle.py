"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker.  The product library never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True,
                   stdout=subprocess.DEVNULL)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.orc_num_cells.argtypes = [vp]
    lib.orc_num_cells.restype = ctypes.c_long
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_set_threads.restype = ctypes.c_int
    lib.orc_constants.argtypes = [vp, vp]
    lib.orc_run_f32.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_float]
    lib.orc_run_f64.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_float]
    lib.orc_probe_f32.argtypes = [vp, ctypes.c_int] + [vp] * 9
    lib.orc_brute_neighbors_f32.argtypes = [ctypes.c_int, vp, ctypes.c_float, vp]
    lib.orc_density_f64.argtypes = [vp, ctypes.c_int, vp, vp]
    lib.orc_vorticity_f64.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    lib.orc_vorticity_force_f64.argtypes = [vp, ctypes.c_int, vp, vp, vp]
    lib.orc_density_grid_f64.argtypes = [vp, ctypes.c_int, vp, vp]
    for f in ("orc_constants", "orc_run_f32", "orc_run_f64", "orc_probe_f32",
              "orc_brute_neighbors_f32", "orc_density_f64", "orc_vorticity_f64",
              "orc_vorticity_force_f64", "orc_density_grid_f64"):
        getattr(lib, f).restype = ctypes.c_int
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _pp(params):
    return ctypes.cast(ctypes.byref(params), ctypes.c_void_p)


def set_threads(n: int) -> int:
    return load().orc_set_threads(n)


def constants(params) -> dict:
    out = np.zeros(19, np.float32)
    assert load().orc_constants(_pp(params), _p(out)) == 0
    keys = ["h", "h2", "poly6", "spiky", "dt", "rho0", "cl_x", "cl_y", "cl_z",
            "min_x", "min_y", "min_z", "max_x", "max_y", "max_z", "n_x", "n_y",
            "n_z", "w_dq"]
    return dict(zip(keys, out.tolist()))


def state_from_particles(parts: np.ndarray, dtype=np.float32):
    """(n, 10) particle records -> float4 pos (x, 0) and vel (v, m), as
    Simulation::init packs them (Simulation.cpp:120-134)."""
    n = parts.shape[0]
    pos = np.zeros((n, 4), dtype)
    vel = np.zeros((n, 4), dtype)
    pos[:, :3] = parts[:, 1:4]
    vel[:, :3] = parts[:, 4:7]
    vel[:, 3] = parts[:, 0]
    return pos, vel


def run_f32(params, pos4, vel4, nsteps: int, wave: float = 0.0):
    pos = np.ascontiguousarray(pos4, np.float32).copy()
    vel = np.ascontiguousarray(vel4, np.float32).copy()
    rc = load().orc_run_f32(_pp(params), pos.shape[0], _p(pos), _p(vel), nsteps, wave)
    assert rc == 0, "oracle rejected the parameters"
    return pos, vel


def run_f64(params, pos4, vel4, nsteps: int, wave: float = 0.0):
    pos = np.ascontiguousarray(pos4, np.float64).copy()
    vel = np.ascontiguousarray(vel4, np.float64).copy()
    rc = load().orc_run_f64(_pp(params), pos.shape[0], _p(pos), _p(vel), nsteps, wave)
    assert rc == 0, "oracle rejected the parameters"
    return pos, vel


def num_cells(params) -> int:
    c = load().orc_num_cells(_pp(params))
    assert c > 0, "oracle rejected the parameters"
    return int(c)


def probe_f32(params, pos4, vel4, ncells=None) -> dict:
    """ncells, if given, must equal the oracle's cell count (checked)."""
    n = pos4.shape[0]
    c = num_cells(params)
    if ncells is not None and ncells != c:
        raise ValueError(f"cell count {ncells} != oracle grid {c}")
    ncells = c
    pos = np.ascontiguousarray(pos4, np.float32)
    vel = np.ascontiguousarray(vel4, np.float32)
    out = {
        "cell_of": np.zeros(n, np.int32),
        "sorted_ids": np.zeros(n, np.int32),
        "cell_start": np.zeros(ncells + 1, np.int32),
        "nbr_count": np.zeros(n, np.int32),
        "rho": np.zeros(n, np.float32),
        "lambda": np.zeros(n, np.float32),
        "pred": np.zeros((n, 4), np.float32),
    }
    rc = load().orc_probe_f32(_pp(params), n, _p(pos), _p(vel), _p(out["cell_of"]),
                              _p(out["sorted_ids"]), _p(out["cell_start"]),
                              _p(out["nbr_count"]), _p(out["rho"]),
                              _p(out["lambda"]), _p(out["pred"]))
    assert rc == 0
    return out


def brute_neighbors(pred4, h2: float) -> np.ndarray:
    pred = np.ascontiguousarray(pred4, np.float32)
    out = np.zeros(pred.shape[0], np.int32)
    assert load().orc_brute_neighbors_f32(pred.shape[0], _p(pred), h2, _p(out)) == 0
    return out


def density_f64(params, pos4) -> np.ndarray:
    pos = np.ascontiguousarray(pos4, np.float64)
    out = np.zeros(pos.shape[0], np.float64)
    assert load().orc_density_f64(_pp(params), pos.shape[0], _p(pos), _p(out)) == 0
    return out


def density_grid_f64(params, pos4) -> np.ndarray:
    """FP64 density at the given positions through the cell lists of the
    parameters' grid (large n)."""
    pos = np.ascontiguousarray(pos4, np.float64)
    out = np.zeros(pos.shape[0], np.float64)
    assert load().orc_density_grid_f64(_pp(params), pos.shape[0], _p(pos), _p(out)) == 0
    return out


def vorticity_f64(params, pos4, vel4) -> np.ndarray:
    pos = np.ascontiguousarray(pos4, np.float64)
    vel = np.ascontiguousarray(vel4, np.float64)
    out = np.zeros((pos.shape[0], 4), np.float64)
    assert load().orc_vorticity_f64(_pp(params), pos.shape[0], _p(pos), _p(vel),
                                    _p(out)) == 0
    return out


def vorticity_force_f64(params, pos4, omega4) -> np.ndarray:
    """Velocity change of the vorticity confinement force (eq. 16) for the
    given positions and vorticities (FP64)."""
    pos = np.ascontiguousarray(pos4, np.float64)
    om = np.ascontiguousarray(omega4, np.float64)
    out = np.zeros((pos.shape[0], 4), np.float64)
    assert load().orc_vorticity_force_f64(_pp(params), pos.shape[0], _p(pos),
                                          _p(om), _p(out)) == 0
    return out
