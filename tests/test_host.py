"""CPU-side checks of the product library: it loads, exports every entry
point include/pbf.h declares, validates parameters, and its scenario readers
behave as the reference's ConfigReader / PartReader (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import pbf
import scenarios

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "pbf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pbf.load_library()
    names = header_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    # the Python binding declares the same set
    assert sorted(pbf.EXPORTED) == names


def test_params_init_defaults():
    p = pbf.default_params(pbf.PHYSICS_REF)
    assert p.iterations == 4 and p.relaxation == 1e4
    assert p.gravity_y == np.float32(-9.81) and p.xsph_c == np.float32(0.01)
    assert p.scorr_k == 0 and p.vorticity_eps == 0
    q = pbf.default_params(pbf.PHYSICS_PAPER)
    assert q.relaxation == 600 and q.scorr_n == 4 and q.scorr_dq == np.float32(0.2)
    assert q.vorticity_eps == np.float32(1e-4)
    with pytest.raises(pbf.PBFError):
        pbf.default_params(7)


def test_read_par_matches_reference_file(golden_dir):
    p, part = pbf.read_par(os.path.join(golden_dir, "two.par"))
    assert part == "two.in"
    assert p.time_step_length == np.float32(0.005)
    assert (p.x_min, p.x_max, p.x_n) == (0.0, 1.0, 100.0)
    assert p.rest_density == np.float32(998200.0)
    assert p.physics == pbf.PHYSICS_REF and p.iterations == 4
    p2, part2 = pbf.read_par(os.path.join(golden_dir, "dam_coarse.par"))
    assert part2 == "dam_coarse.in"
    assert (p2.x_max, p2.x_n, p2.rest_density) == (30.0, 30.0, np.float32(9.261))


def test_read_par_physics_keys_and_unknown_key(tmp_path):
    f = tmp_path / "s.par"
    f.write_text("part_input_file a.in\nx_max 2\nx_n 20\nphysics paper\n"
                 "iterations 3\nscorr_k 0.002\nsphere 1 2 3 0.5\nbogus 1\n"
                 "launch_grid 2048\nscan_cells 4096\n")
    p, part = pbf.read_par(str(f))
    assert part == "a.in" and p.physics == pbf.PHYSICS_PAPER
    assert p.iterations == 3 and p.scorr_k == np.float32(0.002)
    assert p.launch_grid == 2048 and p.scan_cells == 4096
    assert list(p.sphere) == [1.0, 2.0, 3.0, 0.5]
    assert p.x_max == 2.0 and p.x_n == 20.0


def test_read_par_physics_key_order_does_not_matter(tmp_path):
    """Physics keys given before `physics paper` keep their values; the
    mode only supplies the defaults of the keys the file does not set."""
    f = tmp_path / "o.par"
    f.write_text("x_max 2\nx_n 20\niterations 6\nblock_size 64\n"
                 "scorr_n 2\nphysics paper\nxsph_c 0.02\n")
    p, _ = pbf.read_par(str(f))
    assert p.physics == pbf.PHYSICS_PAPER
    assert p.iterations == 6 and p.block_size == 64 and p.scorr_n == 2.0
    assert p.xsph_c == np.float32(0.02)
    # the unset keys take the PAPER defaults
    assert p.relaxation == 600.0 and p.scorr_k == np.float32(0.001)
    assert p.vorticity_eps == np.float32(1e-4)
    g = tmp_path / "bad.par"
    g.write_text("iterations 2\nphysics fast\n")
    with pytest.raises(pbf.PBFError):
        pbf.read_par(str(g))


def test_read_par_missing_file():
    with pytest.raises(pbf.PBFError):
        pbf.read_par("/nonexistent/file.par")


def test_read_particles(golden_dir):
    a = pbf.read_particles(os.path.join(golden_dir, "two.in"))
    assert a.shape == (2, 10)
    assert np.array_equal(a[1, :4], np.float32([1, 0.49577011, 0.5, 0.5]))
    b = pbf.read_particles(os.path.join(golden_dir, "27.in"))
    assert b.shape[0] == 27
    c = pbf.read_particles(os.path.join(golden_dir, "dam.in"))
    ref = np.loadtxt(os.path.join(golden_dir, "dam.in"), skiprows=1)
    assert c.shape[0] == ref.shape[0] == 1000
    assert np.array_equal(c[:, :7], ref.astype(np.float32))


@pytest.mark.parametrize("name", ["two", "two_updown", "27", "box_small", "box",
                                  "box_big", "dam", "dam2", "dam_coarse",
                                  "dam_miles", "drop"])
def test_reference_scenarios_load(golden_dir, name):
    """Every scenario file the reference ships with a particle file loads
    unchanged: the .par names its particle file, and the particles equal
    the file's numeric columns (m, x, v) row by row."""
    p, part = pbf.read_par(os.path.join(golden_dir, name + ".par"))
    assert part == name + ".in"
    assert p.physics == pbf.PHYSICS_REF and p.time_step_length > 0
    a = pbf.read_particles(os.path.join(golden_dir, part))
    ref = np.loadtxt(os.path.join(golden_dir, part), skiprows=1, ndmin=2)
    assert a.shape[0] == ref.shape[0] > 0
    assert np.array_equal(a[:, :7], ref[:, :7].astype(np.float32))


def test_read_particles_stops_at_bad_mass(tmp_path):
    f = tmp_path / "bad.in"
    f.write_text("4\n1 0 0 0 0 0 0\n1 1 1 1 0 0 0\nxyz 1 1 1\n1 2 2 2 0 0 0\n")
    a = pbf.read_particles(str(f))
    assert a.shape[0] == 2  # PartReader.cpp:56-60: break on a bad mass


def test_read_particles_short_file(tmp_path):
    f = tmp_path / "short.in"
    f.write_text("5\n1 0 0 0 0 0 0\n")
    assert pbf.read_particles(str(f)).shape[0] == 1


def test_create_rejects_bad_params():
    p = pbf.default_params(pbf.PHYSICS_REF)
    p.x_max = p.y_max = p.z_max = 1.0
    p.x_n = p.y_n = p.z_n = 10.0
    p.time_step_length, p.rest_density = 0.01, 1000.0
    parts = np.zeros((1, 10), np.float32)
    bad = p.copy()
    bad.x_n = 10.5
    with pytest.raises(pbf.PBFError) as e:
        pbf.Simulation(bad, parts)
    assert e.value.code == pbf.ERR_INVALID
    bad = p.copy()
    bad.iterations = 0
    with pytest.raises(pbf.PBFError):
        pbf.Simulation(bad, parts)
    bad = p.copy()
    bad.x_max = bad.x_min
    with pytest.raises(pbf.PBFError):
        pbf.Simulation(bad, parts)
    # 1000^3 cells fit 2^30 for REF, not once PAPER splits x in 4 (the
    # product overflowed a 32-bit cell count before it was checked)
    big = pbf.default_params(pbf.PHYSICS_PAPER)
    for name in ("time_step_length", "rest_density"):
        setattr(big, name, getattr(p, name))
    big.x_max = big.y_max = big.z_max = 100.0
    big.x_n = big.y_n = big.z_n = 1000.0
    with pytest.raises(pbf.PBFError) as e:
        pbf.Simulation(big, parts)
    assert e.value.code == pbf.ERR_INVALID and "too many cells" in str(e.value)
    # mixed precision: PAPER physics only, 1 <= fp32_iterations <= iterations
    bad = p.copy()
    bad.precision, bad.fp32_iterations = pbf.PRECISION_MIXED, 1
    with pytest.raises(pbf.PBFError) as e:
        pbf.Simulation(bad, parts)
    assert "PAPER" in str(e.value) or "paper" in str(e.value)
    q = pbf.default_params(pbf.PHYSICS_PAPER)
    for name in ("x_max", "y_max", "z_max", "x_n", "y_n", "z_n",
                 "time_step_length", "rest_density"):
        setattr(q, name, getattr(p, name))
    q.precision = pbf.PRECISION_MIXED
    for fp32_it in (0, q.iterations + 1):
        bad = q.copy()
        bad.fp32_iterations = fp32_it
        with pytest.raises(pbf.PBFError) as e:
            pbf.Simulation(bad, parts)
        assert "fp32_iterations" in str(e.value)
    # a persistent grid is a multiple of 8 work-groups (one XCD chunk each)
    for grid in (-8, 12, (1 << 20) + 8):
        bad = p.copy()
        bad.launch_grid = grid
        with pytest.raises(pbf.PBFError) as e:
            pbf.Simulation(bad, parts)
        assert e.value.code == pbf.ERR_INVALID and "launch_grid" in str(e.value)
    for cells in (-1, 512, 3072, 8192):
        bad = p.copy()
        bad.scan_cells = cells
        with pytest.raises(pbf.PBFError) as e:
            pbf.Simulation(bad, parts)
        assert e.value.code == pbf.ERR_INVALID and "scan_cells" in str(e.value)


def test_scenario_configs_shapes():
    c0 = scenarios.config(0)
    assert c0.n == 32768 and c0.steps == 100
    p = c0.params[0]
    assert (p.x_max, p.y_max, p.z_max) == (4.0, 2.0, 2.0)
    assert (p.x_n, p.y_n, p.z_n) == (40.0, 20.0, 20.0)
    x = c0.particles[0][:, 1:4]
    assert x.min() >= 0.1 and x[:, 0].max() < 4.0 - 0.1 and x[:, 1].max() < 2.0 - 0.1
    # deterministic
    assert np.array_equal(scenarios.config(0).particles[0], c0.particles[0])
    # jitter bounded by 0.01 h
    k = np.arange(32)
    site = scenarios.MARGIN + k * scenarios.SPACING
    assert np.abs(x[:32, 2] - site).max() <= 0.001 + 1e-7


def test_batch_config_statistics():
    c3 = scenarios.batch_config()
    counts = np.array(c3.meta["counts"])
    assert len(counts) == 256
    assert counts.min() >= 2048 and counts.max() <= 65536
    assert 3.0e6 < counts.sum() < 7.0e6
    for p, parts in zip(c3.params, c3.particles):
        box = np.array([p.x_max, p.y_max, p.z_max])
        x = parts[:, 1:4]
        assert x.min() > 0.1 - 0.002 and np.all(x.max(0) < box - 0.1 + 0.002)
        cells = np.array([p.x_n, p.y_n, p.z_n])
        assert np.allclose(box / cells, 0.1, rtol=1e-6)


def test_splitmix64_known_values():
    # splitmix64 of seed 0: first outputs (Steele, Lea & Flood 2014 reference
    # implementation, as used by xoshiro seeding)
    v = scenarios.splitmix64(0, 3)
    assert [int(x) for x in v] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                   0x06C45D188009454F]


@settings(max_examples=25, deadline=None, derandomize=True)
@given(n=st.integers(0, 300), seed=st.integers(0, 2**31 - 1),
       box=st.tuples(*[st.integers(1, 400)] * 3))
def test_scenario_files_round_trip(tmp_path_factory, n, seed, box):
    """Seeded property: particles and box parameters written in the
    reference's .in / .par formats read back to the same float32 values."""
    d = tmp_path_factory.mktemp("rt")
    rng = np.random.default_rng(seed)
    parts = np.zeros((n, 10), np.float32)
    parts[:, :7] = rng.normal(0.0, 3.0, (n, 7)).astype(np.float32)
    parts[:, 0] = np.abs(parts[:, 0]) + np.float32(0.5)   # masses > 0
    scenarios.write_in(str(d / "p.in"), parts)
    got = pbf.read_particles(str(d / "p.in"))
    assert got.shape == (n, 10)
    assert np.array_equal(got[:, :7], parts[:, :7])
    p = scenarios.make_params(tuple(0.1 * b for b in box), pbf.PHYSICS_REF,
                              rest_density=float(rng.uniform(100, 9000)))
    scenarios.write_par(str(d / "s.par"), p, "p.in")
    q, part_file = pbf.read_par(str(d / "s.par"))
    assert part_file == "p.in"
    for k in ("time_step_length", "time_end", "x_min", "x_max", "y_min",
              "y_max", "z_min", "z_max", "x_n", "y_n", "z_n", "rest_density"):
        assert np.float32(getattr(q, k)) == np.float32(getattr(p, k)), k


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it says so and fails
    before touching a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_wave_generator_sequence_matches_reference_formula(tmp_path):
    """The Runner's moving-wall offset (csrc/WaveGenerator.hpp), compiled
    with g++, against a numpy restatement of src/Runner.cpp:91-104: push
    length (x_max - x_min) / 3.0f and phase -M_PI / 2.0f stored as cl_float,
    sin(2.0f * M_PI * frequency * wave + phase) evaluated in double, the
    result stored as cl_float, wave accumulated in float by dt."""
    src = tmp_path / "wave.cpp"
    src.write_text(
        '#include <cstdio>\n#include "WaveGenerator.hpp"\n'
        "int main() {\n"
        "  const float xmin = 0.0f, xmax = 4.0f, dt = 1.0f / 120.0f;\n"
        "  const float push = (xmax - xmin) / 3.0f;\n"
        "  float wave = 0.0f;\n"
        "  for (int k = 0; k < 400; ++k) {\n"
        "    std::printf(\"%a\\n\", (double)waveGeneratorValue(push, wave));\n"
        "    wave += dt;\n"
        "  }\n}\n")
    exe = tmp_path / "wave"
    import subprocess
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "position-based-fluids_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    got = [float.fromhex(v) for v in
           subprocess.run([str(exe)], check=True, capture_output=True,
                          text=True).stdout.split()]
    f32 = np.float32
    push = f32(f32(4.0) - f32(0.0)) / f32(3.0)
    phase = float(f32(-np.pi / 2.0))

    def sequence(ph):
        wave, out = f32(0.0), []
        for _ in range(400):
            arg = 2.0 * np.pi * float(f32(1.0)) * float(wave) + ph
            v = np.sin(arg) * float(push) / 2.0 + float(push / f32(2.0))
            out.append(float(f32(v)))
            wave = f32(wave + f32(1.0 / 120.0))
        return out

    expect = sequence(phase)
    assert got == expect
    # the exact-double phase gives other floats: the test tells them apart
    assert sum(a != b for a, b in zip(sequence(-np.pi / 2.0), expect)) > 50


def test_bench_algorithmic_bytes():
    """The roofline's byte model: 324 + 60 I bytes per particle-step (PAPER),
    244 + 60 I (REF), 16 per grid cell; 570.9 B at the 1M-particle dam break
    (432,000 cells); a delta pass 36 B, a lambda pass 24 B per particle."""
    import bench
    assert bench.step_bytes_per_particle(4, pbf.PHYSICS_PAPER) == 564.0
    assert bench.step_bytes_per_particle(4, pbf.PHYSICS_REF) == 484.0
    assert bench.step_bytes_per_particle(1, pbf.PHYSICS_PAPER) == 384.0
    c1 = scenarios.config(1)
    p = c1.params[0]
    cells = p.x_n * p.y_n * p.z_n
    assert cells == 432000
    b = bench.step_bytes_per_particle(4, pbf.PHYSICS_PAPER) + 16.0 * cells / c1.n
    assert b == pytest.approx(570.912)
    assert bench.pass_bytes("delta") == 36.0 and bench.pass_bytes("lambda") == 24.0
    info = bench.cpu_info()
    assert info["nproc"] >= 1 and info["usable"] >= 1 and info["cpu_model"]
