"""ctypes binding of libpbf.so (the C ABI declared in include/pbf.h).

Mirrors the reference's solver interface for the per-step path:
``Simulation`` (src/Simulation.hpp:30-197) with ``step``, ``dump_data``
(dumpData), ``set_wave_generator``, ``get_sizes_min/max`` and
``get_number_particles``.  Construction covers the reference's
``Simulation(...)`` + ``init()`` + ``initCells()`` (src/Simulation.cpp:34-230).

The library is loaded from this directory.  If it is missing the import of
this module still works, but creating a simulation raises ``PBFError``:
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbf.so")

# kernel classes in the library's order (pbf_profile_steps, pbf_get_launch)
KERNEL_CLASSES = ("predict_count", "scan", "scatter", "reorder", "lambda",
                  "delta", "xsph", "vorticity", "search")

PHYSICS_REF = 0
PHYSICS_PAPER = 1
PRECISION_FP32 = 0
PRECISION_MIXED = 1

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_STATE = 3
ERR_NUMERIC = 4


class Params(ctypes.Structure):
    """pbf_params: ConfigParameters (src/Parameters.hpp:12-31) + physics."""

    _fields_ = [
        ("time_step_length", ctypes.c_float),
        ("time_end", ctypes.c_float),
        ("x_min", ctypes.c_float),
        ("x_max", ctypes.c_float),
        ("y_min", ctypes.c_float),
        ("y_max", ctypes.c_float),
        ("z_min", ctypes.c_float),
        ("z_max", ctypes.c_float),
        ("x_n", ctypes.c_float),
        ("y_n", ctypes.c_float),
        ("z_n", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("physics", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("relaxation", ctypes.c_float),
        ("scorr_k", ctypes.c_float),
        ("scorr_n", ctypes.c_float),
        ("scorr_dq", ctypes.c_float),
        ("xsph_c", ctypes.c_float),
        ("vorticity_eps", ctypes.c_float),
        ("gravity_y", ctypes.c_float),
        ("precision", ctypes.c_int32),
        ("fp32_iterations", ctypes.c_int32),
        ("sphere", ctypes.c_float * 4),
        ("block_size", ctypes.c_int32),
        ("launch_grid", ctypes.c_int32),
        ("scan_cells", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 5),
    ]

    def copy(self) -> "Params":
        p = Params()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self), ctypes.sizeof(Params))
        return p

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class Particle(ctypes.Structure):
    """pbf_particle: layout of the reference Particle (src/Particle.hpp:9-14)."""

    _fields_ = [
        ("m", ctypes.c_float),
        ("x", ctypes.c_float * 3),
        ("v", ctypes.c_float * 3),
        ("a", ctypes.c_float * 3),
    ]


assert ctypes.sizeof(Particle) == 40


class PBFError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"pbf error {code}: {message}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every exported entry point of include/pbf.h
_P = ctypes.POINTER
_SIGNATURES = [
    ("pbf_params_init", ctypes.c_int, [_P(Params), ctypes.c_int32]),
    ("pbf_create", ctypes.c_int,
     [_P(Params), ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_void_p)]),
    ("pbf_batch_create", ctypes.c_int,
     [ctypes.c_uint32, _P(Params), _P(ctypes.c_void_p), _P(ctypes.c_uint32),
      _P(ctypes.c_void_p)]),
    ("pbf_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    ("pbf_synchronize", ctypes.c_int, [ctypes.c_void_p]),
    ("pbf_dump", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("pbf_batch_dump", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    ("pbf_positions_to_device", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]),
    ("pbf_set_wave_generator", ctypes.c_int, [ctypes.c_void_p, ctypes.c_float]),
    ("pbf_get_box", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_float), _P(ctypes.c_float)]),
    ("pbf_num_particles", ctypes.c_uint32, [ctypes.c_void_p]),
    ("pbf_num_scenes", ctypes.c_uint32, [ctypes.c_void_p]),
    ("pbf_num_cells", ctypes.c_uint32, [ctypes.c_void_p]),
    ("pbf_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("pbf_last_error", ctypes.c_char_p, []),
    ("pbf_probe", ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 7),
    ("pbf_diagnostics", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double)]),
    ("pbf_profile_steps", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_double), _P(ctypes.c_uint32),
      _P(ctypes.c_char_p), _P(ctypes.c_uint32)]),
    ("pbf_time_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_double)]),
    ("pbf_time_kernel", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, _P(ctypes.c_double)]),
    ("pbf_get_config", ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32)]),
    ("pbf_get_launch", ctypes.c_int,
     [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_int32), ctypes.c_int32,
      _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    ("pbf_calibrate_copy", ctypes.c_int,
     [ctypes.c_uint64, ctypes.c_uint32, _P(ctypes.c_double)]),
    ("pbf_step_noalloc", ctypes.c_int,
     [_P(Params), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
      ctypes.c_uint32, ctypes.c_void_p]),
    ("pbf_read_par", ctypes.c_int,
     [ctypes.c_char_p, _P(Params), ctypes.c_char_p, ctypes.c_size_t]),
    ("pbf_read_particles", ctypes.c_int,
     [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint32, _P(ctypes.c_uint32)]),
]

EXPORTED = [name for name, _, _ in _SIGNATURES]


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libpbf.so once and declare every signature of include/pbf.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PBFError(ERR_STATE, f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(code: int) -> None:
    if code != OK:
        msg = _lib.pbf_last_error().decode(errors="replace")
        raise PBFError(code, msg)


def default_params(physics: int = PHYSICS_PAPER) -> Params:
    lib = load_library()
    p = Params()
    _check(lib.pbf_params_init(ctypes.byref(p), physics))
    return p


def particles_to_array(particles) -> np.ndarray:
    """Accept an (n, 10) / (n, 7) float array [m, x, y, z, vx, vy, vz(, a)] or a
    sequence of Particle and return a C-contiguous float32 (n, 10) array."""
    if isinstance(particles, np.ndarray):
        a = np.asarray(particles, dtype=np.float32)
        if a.ndim != 2 or a.shape[1] not in (7, 10):
            raise ValueError("particles must have shape (n, 7) or (n, 10)")
        if a.shape[1] == 7:
            a = np.concatenate([a, np.zeros((a.shape[0], 3), np.float32)], axis=1)
        return np.ascontiguousarray(a)
    out = np.zeros((len(particles), 10), np.float32)
    for k, q in enumerate(particles):
        out[k] = [q.m, *q.x, *q.v, *q.a]
    return out


def pack_state(pos: np.ndarray, vel: np.ndarray) -> np.ndarray:
    """float4 pos (x, y, z, *) and vel (vx, vy, vz, m) -> (n, 10) particles."""
    n = pos.shape[0]
    out = np.zeros((n, 10), np.float32)
    out[:, 0] = vel[:, 3]
    out[:, 1:4] = pos[:, :3]
    out[:, 4:7] = vel[:, :3]
    return out


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Simulation:
    """One scene on the GPU (src/Simulation.hpp:30-197)."""

    def __init__(self, params: Params, particles):
        lib = load_library()
        self._lib = lib
        self.params = params.copy()
        self._parts = particles_to_array(particles)
        self.n = self._parts.shape[0]
        h = ctypes.c_void_p()
        _check(lib.pbf_create(ctypes.byref(self.params), _ptr(self._parts),
                              self.n, ctypes.byref(h)))
        self._h = h

    # -- reference API ------------------------------------------------------
    def init(self) -> None:
        """Simulation::init (device state is set up by the constructor)."""

    def init_cells(self) -> None:
        """Simulation::initCells: cell tables are rebuilt every step."""

    def step(self, nsteps: int = 1) -> None:
        _check(self._lib.pbf_step(self._h, nsteps))

    def synchronize(self) -> None:
        _check(self._lib.pbf_synchronize(self._h))

    def dump_data(self):
        """Simulation::dumpData: (pos (n,4), vel (n,4)) in input order."""
        pos = np.zeros((self.n, 4), np.float32)
        vel = np.zeros((self.n, 4), np.float32)
        _check(self._lib.pbf_dump(self._h, _ptr(pos), _ptr(vel)))
        return pos, vel

    dump = dump_data

    def set_wave_generator(self, value: float) -> None:
        _check(self._lib.pbf_set_wave_generator(self._h, value))

    def positions_to_device(self, dst: int, scene: int = 0) -> None:
        """(x, |v|) of every particle of `scene`, float4 by particle index,
        written into the device buffer at address `dst` (the positions
        buffer a viewer draws; a mapped GL buffer or hipMalloc memory)."""
        _check(self._lib.pbf_positions_to_device(self._h, scene,
                                                 ctypes.c_void_p(dst)))

    def get_sizes_min(self):
        mn = (ctypes.c_float * 4)()
        mx = (ctypes.c_float * 4)()
        _check(self._lib.pbf_get_box(self._h, mn, mx))
        return tuple(mn)

    def get_sizes_max(self):
        mn = (ctypes.c_float * 4)()
        mx = (ctypes.c_float * 4)()
        _check(self._lib.pbf_get_box(self._h, mn, mx))
        return tuple(mx)

    def get_number_particles(self) -> int:
        return int(self._lib.pbf_num_particles(self._h))

    # -- parity hooks and measurement ---------------------------------------
    def num_cells(self) -> int:
        return int(self._lib.pbf_num_cells(self._h))

    def probe(self) -> dict:
        n, c = self.n, self.num_cells()
        out = {
            "cell_of": np.zeros(n, np.int32),
            "sorted_ids": np.zeros(n, np.int32),
            "cell_start": np.zeros(c + 1, np.int32),
            "nbr_count": np.zeros(n, np.int32),
            "rho": np.zeros(n, np.float32),
            "lambda": np.zeros(n, np.float32),
            "pred": np.zeros((n, 4), np.float32),
        }
        _check(self._lib.pbf_probe(
            self._h, _ptr(out["cell_of"]), _ptr(out["sorted_ids"]),
            _ptr(out["cell_start"]), _ptr(out["nbr_count"]), _ptr(out["rho"]),
            _ptr(out["lambda"]), _ptr(out["pred"])))
        return out

    def diagnostics(self) -> dict:
        d = (ctypes.c_double * 8)()
        _check(self._lib.pbf_diagnostics(self._h, d))
        return {"max_compression": d[0], "mean_abs_density_err": d[1],
                "momentum": (d[2], d[3], d[4]), "kinetic_energy": d[5],
                "count": d[6], "max_abs_density_err": d[7]}

    def profile(self, nsteps: int) -> dict:
        cap = 16
        ms = (ctypes.c_double * cap)()
        launches = (ctypes.c_uint32 * cap)()
        names = (ctypes.c_char_p * cap)()
        ncls = ctypes.c_uint32(cap)
        _check(self._lib.pbf_profile_steps(self._h, nsteps, ms, launches, names,
                                           ctypes.byref(ncls)))
        return {names[k].decode(): (ms[k], launches[k]) for k in range(ncls.value)}

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        _check(self._lib.pbf_time_steps(self._h, nsteps, ctypes.byref(ms)))
        return ms.value

    def time_kernel(self, cls: str, reps: int = 20) -> float:
        """Average ms of one launch of kernel class `cls` (search, lambda,
        delta), timed over `reps` back-to-back launches in a graph."""
        ms = ctypes.c_double()
        _check(self._lib.pbf_time_kernel(self._h, cls.encode(), reps,
                                         ctypes.byref(ms)))
        return ms.value

    def config(self):
        cfg = (ctypes.c_int32 * 4)()
        _check(self._lib.pbf_get_config(self._h, cfg))
        return tuple(cfg)

    def launch(self) -> dict:
        """Launch form per kernel class: {class: (work-group size, grid)},
        grid 0 for one work-group per tile, -1 for a persistent launch of
        the resident work-groups, else the work-group count of a persistent
        launch; and "scan_cells": cells per work-group of the cell-count
        scan."""
        block = (ctypes.c_int32 * 16)()
        grid = (ctypes.c_int32 * 16)()
        count = ctypes.c_int32()
        cells = ctypes.c_int32()
        _check(self._lib.pbf_get_launch(self._h, block, grid, 16,
                                        ctypes.byref(count), ctypes.byref(cells)))
        out = {KERNEL_CLASSES[c]: (block[c], grid[c]) for c in range(count.value)}
        out["scan_cells"] = cells.value
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            _check(self._lib.pbf_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSimulation(Simulation):
    """Many independent scenes advanced together (one launch per kernel)."""

    def __init__(self, params: Sequence[Params], particles: Sequence):
        lib = load_library()
        self._lib = lib
        if len(params) != len(particles) or not params:
            raise ValueError("need one Params per particle set")
        self.params_list = [p.copy() for p in params]
        self.params = self.params_list[0]
        self._parts = [particles_to_array(p) for p in particles]
        ns = len(params)
        arr = (Params * ns)(*self.params_list)
        ptrs = (ctypes.c_void_p * ns)(*[a.ctypes.data for a in self._parts])
        counts = (ctypes.c_uint32 * ns)(*[a.shape[0] for a in self._parts])
        self.counts = [a.shape[0] for a in self._parts]
        self.n = sum(self.counts)
        h = ctypes.c_void_p()
        _check(lib.pbf_batch_create(ns, arr, ptrs, counts, ctypes.byref(h)))
        self._h = h

    def dump_scene(self, scene: int):
        n = self.counts[scene]
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        _check(self._lib.pbf_batch_dump(self._h, scene, _ptr(pos), _ptr(vel)))
        return pos, vel

    def dump_data(self):
        raise PBFError(ERR_STATE, "use dump_scene on a batched simulation")

    dump = dump_data

    def num_scenes(self) -> int:
        return int(self._lib.pbf_num_scenes(self._h))


def calibrate_copy(nbytes: int, reps: int = 10) -> float:
    """Average time (ms) of a plain float4 copy of nbytes (read + write)."""
    lib = load_library()
    ms = ctypes.c_double()
    _check(lib.pbf_calibrate_copy(nbytes, reps, ctypes.byref(ms)))
    return ms.value


def read_par(path: str):
    """ConfigReader::read (src/io/ConfigReader.cpp:19-94) -> (Params, part file)."""
    lib = load_library()
    p = Params()
    buf = ctypes.create_string_buffer(4096)
    _check(lib.pbf_read_par(path.encode(), ctypes.byref(p), buf, len(buf)))
    return p, buf.value.decode()


def read_particles(path: str) -> np.ndarray:
    """PartReader::read (src/io/PartReader.cpp:21-97) -> (n, 10) float32."""
    lib = load_library()
    n = ctypes.c_uint32(0)
    _check(lib.pbf_read_particles(path.encode(), None, 0, ctypes.byref(n)))
    out = np.zeros((n.value, 10), np.float32)
    cap = n.value
    _check(lib.pbf_read_particles(path.encode(), _ptr(out), cap, ctypes.byref(n)))
    return out[: n.value]
